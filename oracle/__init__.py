"""FP64 host oracle for the ALS update path — test infrastructure only."""
