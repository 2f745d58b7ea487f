"""Host-side Python access to the MI355X correlator-node data path.

job      -- the Correlation_parameters / delay model / pulsar description
synth    -- seeded synthetic station data and the input-node HEADER stream
configs  -- builders for the synthetic correlation jobs
capi     -- ctypes binding of include/sfxg.h (libsfxg.so)
"""
