"""MI355X HIP backend for the portable Digital ICE processing core."""
