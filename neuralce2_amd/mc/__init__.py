"""Monte Carlo entry points (the neuralce-mc command line)."""
