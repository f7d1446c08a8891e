"""CPU oracle package (test infrastructure; see oracle/blast_oracle.cpp)."""
