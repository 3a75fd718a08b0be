"""CPU oracle (test infrastructure only): FP64 restatement of the reference's
pressure path for checking the HIP engine and for the benchmark's CPU baseline."""
