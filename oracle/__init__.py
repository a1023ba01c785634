"""Oracle package: test infrastructure only (see refcpu.c header)."""
