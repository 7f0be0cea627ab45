"""TEST INFRASTRUCTURE ONLY: FP64 CPU oracle of the element assembly (see oracle/oracle.cpp)."""
