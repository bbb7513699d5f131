"""CPU oracle (test infrastructure): ctypes binding of liboracle.so."""
