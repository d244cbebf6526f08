"""Benchmark harness of the state-update path (not part of the shipped library).

bench.py at the repository root drives these modules:
  workloads   synthetic inputs of the five benchmark configurations (seeded, host-generated)
  dm          14-qubit density-matrix layers, both layouts, per-kernel GB/s and CPU baseline
  qaoa        4096-instance batched QAOA, evaluations/s and CPU baseline
  screening   learned screening for QAOA angle search (surrogate training and search)
  cpu         CPU-baseline method (two-level timing loop) and host description (lscpu)

Only the CPU-baseline legs bind the CPU libraries under oracle/ (the compiled reference or the
restatement), as the baseline being timed; the GPU figures come from orkan_amd/lib alone.
"""
