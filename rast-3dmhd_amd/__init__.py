"""MI355X-native explicit compressible-MHD time step (Rast 3dmhd STEP)."""
