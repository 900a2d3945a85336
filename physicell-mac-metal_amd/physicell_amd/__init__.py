"""MI355X (gfx950) step hot path of a PhysiCell-style simulator.

hip        -- ctypes binding of libphysicell_hip.so (include/physicell_hip.h)
scenarios  -- seeded synthetic inputs (configs 0-2 and small test cases)
parity     -- snapshot diffs (L-inf / RMSE / relative, bit-exact integer modes)
"""
