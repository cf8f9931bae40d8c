"""CPU oracle of the CCCS dose path -- test infrastructure only (see cccs_oracle.h)."""
