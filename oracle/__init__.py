"""FP64 CPU oracle of the curved-hex ray-tracing method (test infrastructure only)."""
