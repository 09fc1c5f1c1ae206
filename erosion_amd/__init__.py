"""erosion_amd: MI355X-native per-step Stokes solve for the Erosion package."""
