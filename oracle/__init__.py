"""CPU checker of the GPU path tracer (test infrastructure only)."""
