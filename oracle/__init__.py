"""TEST INFRASTRUCTURE ONLY: CPU oracle of the PLENS DGSEM hot path (see oracle.h)."""
