"""FP64 CPU oracle package -- test infrastructure only (see lbm_oracle.h)."""
