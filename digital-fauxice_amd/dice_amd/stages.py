"""Per-stage calibration tables handed to the C ABI.

The streaming reference resolves three stage calibrations per output row
(source stage y+1, score stage y+2, writer stage y+7;
``stage_parameters.py:207-232``) and stages 1..6 during the hidden startup
(``startup.py:231-321``).  ``StageTable`` holds one entry per stage
1..H+6.  ``constant_stage_table`` is the ``stage_parameter_provider=None``
case of ``derive_row_parameter_table`` (``cuda_backend/rowparams.py:60-139``):
every stage carries the caller's fixed parameters.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._abi import DiceStageTable, f32p, i32p, ptr, u8p
from .parameters import AuxiliaryParameters, ReconstructionParameters, ScoreParameters


@dataclass
class StageTable:
    aux_alpha: np.ndarray
    aux_offset: np.ndarray
    aux_replacement: np.ndarray
    aux_has_replacement: np.ndarray
    score_base: np.ndarray
    coarse_reference: np.ndarray
    gate_secondary: np.ndarray
    row_gate: np.ndarray

    @classmethod
    def empty(cls, count: int) -> "StageTable":
        return cls(
            aux_alpha=np.zeros(count, np.float32),
            aux_offset=np.zeros(count, np.float32),
            aux_replacement=np.zeros(count, np.float32),
            aux_has_replacement=np.zeros(count, np.uint8),
            score_base=np.zeros(count, np.float32),
            coarse_reference=np.zeros(count, np.float32),
            gate_secondary=np.zeros(count, np.uint8),
            row_gate=np.zeros(count, np.int32),
        )

    @property
    def count(self) -> int:
        return int(self.aux_alpha.shape[0])

    def as_struct(self) -> DiceStageTable:
        table = DiceStageTable()
        table.count = self.count
        table.aux_alpha = ptr(self.aux_alpha, f32p._type_)
        table.aux_offset = ptr(self.aux_offset, f32p._type_)
        table.aux_replacement = ptr(self.aux_replacement, f32p._type_)
        table.aux_has_replacement = ptr(self.aux_has_replacement, u8p._type_)
        table.score_base = ptr(self.score_base, f32p._type_)
        table.coarse_reference = ptr(self.coarse_reference, f32p._type_)
        table.gate_secondary = ptr(self.gate_secondary, u8p._type_)
        table.row_gate = ptr(self.row_gate, i32p._type_)
        return table


def constant_stage_table(
    height: int,
    *,
    auxiliary_parameters: AuxiliaryParameters,
    score_parameters: ScoreParameters,
    reconstruction_parameters: ReconstructionParameters,
) -> StageTable:
    count = height + 6
    table = StageTable.empty(count)
    table.aux_alpha[:] = np.float32(auxiliary_parameters.alpha)
    table.aux_offset[:] = np.float32(auxiliary_parameters.calibration_offset)
    if auxiliary_parameters.alpha_one_replacement is not None:
        table.aux_replacement[:] = np.float32(auxiliary_parameters.alpha_one_replacement)
        table.aux_has_replacement[:] = 1
    table.score_base[:] = np.float32(score_parameters.base_primary)
    table.coarse_reference[:] = np.float32(reconstruction_parameters.coarse_reference)
    table.gate_secondary[:] = int(bool(reconstruction_parameters.driver_gate_secondary))
    table.row_gate[:] = int(reconstruction_parameters.row_reconstruction_gate)
    return table
