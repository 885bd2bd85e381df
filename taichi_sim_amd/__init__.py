"""MI355X-native staggered-grid velocity-stress time loop (elastic / Zener, CPML).

* ``taichi_sim_amd.sim_support`` -- host setup with the reference's API
  (Simulator, Simulator3D, SimulationROI, probes, EmissionLaw ...).
* ``taichi_sim_amd.backend`` -- drop-in ``implementation()`` backends
  (SimulatorHIP, SimulatorHIPTaichiElastic, SimulatorHIP3D).
* ``taichi_sim_amd.engine.VSESim`` -- one device-resident simulation on the
  libvse C-ABI (include/vse.h).
"""
__version__ = "0.1.0"
