"""femorph_amd: MI355X-native explicit stabilized fluid time step (feMorph's
femVMSExplicitFluidSolver / femSUPGExplicitFluidSolver) behind a C ABI."""
