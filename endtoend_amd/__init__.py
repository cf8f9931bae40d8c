"""MI355X-native CCCS dose engine (dosecalc-beam / dosecalc-beamlet hot path) -- Python bindings."""
