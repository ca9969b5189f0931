"""karam_amd — MI355X-native Newton-Raphson inner loop of the AcuteSim engine."""
