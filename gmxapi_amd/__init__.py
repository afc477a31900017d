"""MI355X-native NBNxM short-range nonbonded engine (LJ + Ewald real space)."""
