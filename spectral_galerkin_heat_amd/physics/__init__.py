"""Physics kernels of the MI355X backend."""
