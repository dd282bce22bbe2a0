"""MI355X-native river-reach lattice-Boltzmann solver.

The 3-D reach solver of the boat tests (D3Q19 regularized LBM with
Smagorinsky LES, wall model, actuator disks, inlet/outlet rows, rigid lid)
runs as hand-written HIP kernels for gfx950 in libwfs.so, driven from Python
through a ctypes C ABI (include/wfs.h). Module layout mirrors the reference
package's solver path:

    lbm3d        LBM3D, the solver object
    reach3d      case builder, Run3D driver, profiles/compare readouts
    boat_tests   the boat-test sweeps
    geometry     Section
    hydraulics1d discharge for a level (q_ref)
"""
__version__ = "0.1.0"
