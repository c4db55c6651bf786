"""CW k-space PSTD engine (linear and Westervelt) running in libcaustica_hip."""
