"""C++/OpenMP restatement of the reference algorithm for the CPU baseline (test
infrastructure: used by the benchmark's CPU-baseline leg and by tests only)."""
