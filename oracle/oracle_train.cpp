// Per-sample e-prop oracle (filled in with the trainer).
