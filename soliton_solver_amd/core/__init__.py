"""Theory-independent core: parameters, the Simulation driver, native kernel handles."""
