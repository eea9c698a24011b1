"""Particles: the species container of chimeraCL (chimeraCL/particles.py:9-114),
with the same constructor, Args derivation and methods; the device work is
done by ParticleMethodsHIP."""
from __future__ import annotations

import numpy as np

from .methods.generic_methods_hip import LazyArgs, default_comm
from .methods.particles_methods_hip import ParticleMethodsHIP

# weight -> charge in pC per micron of wavelength (particles.py:96), from the
# CODATA 2018 values of scipy.constants: 4 pi^2 m_e c^2 eps0 1e6 / e
_M_E, _C, _E, _EPS0 = 9.1093837015e-31, 299792458.0, 1.602176634e-19, 8.8541878128e-12
_W2PC = 4 * np.pi ** 2 * _M_E * _C ** 2 * _EPS0 * 1e6 / _E

# configuration defaults of a species (particles.py:55-70)
_ARG_DEFAULTS = (("dt", 1.0), ("dx", 1.0), ("dr", 1.0), ("charge", -1.0), ("mass", 1.0),
                 ("dens", 1.0))


class Particles(ParticleMethodsHIP):
    def __init__(self, configs_in, comm=None):
        self.import_comm(comm if comm is not None else default_comm())
        self.set_global_working_group_size()
        self.DataDev = {}
        self.init_particle_methods()
        self._process_configs(configs_in)
        self.send_args_to_dev()
        self._init_data_on_dev()

    def sort_parts(self, grid):
        """particles.py:23-29."""
        if self.Args["Np"] == 0:
            self.flag_sorted = True
        if not self.flag_sorted:
            self.index_sort(grid)
            self.flag_sorted = True

    def add_particles(self, domain_in=None, beam_in=None, source=None):
        """particles.py:31-40."""
        if source is None:
            if domain_in is not None:
                self.make_new_domain(domain_in)
            elif beam_in is not None:
                self.make_new_beam(beam_in)
            self.add_new_particles()
        else:
            self.add_new_particles(source=source)

    def align_parts(self):
        """particles.py:42-51."""
        if self.Args["Np"] == 0:
            return
        self.align_and_damp(comps_align=self._arg_names())

    def _process_configs(self, configs_in):
        """particles.py:53-100."""
        a = self.Args = LazyArgs(configs_in)
        a["Np"] = a["Np_stay"] = 0
        for key, default in _ARG_DEFAULTS:
            a.setdefault(key, default)
        a["dt_2"] = 0.5 * a["dt"]
        # macro-particle weight of a cell-filled plasma: the charge of one
        # cell of density dens, 2 pi dx dr (times r, applied per particle),
        # shared among the prod(Nppc) particles of the cell
        if "Nppc" in a:
            a["Nppc"] = np.array(a["Nppc"], dtype=np.uint32)
            a["w0"] = 2 * np.pi * a["dx"] * a["dr"] * a["dens"] / np.prod(a["Nppc"])
            a["ddx"] = a["dx"] / a["Nppc"][0]
        else:
            a["ddx"] = 1.0
        # Boris factor 2 pi dt q/m (time in lambda0/c, fields in m c omega0/e)
        a["FactorPush"] = 2 * np.pi * a["dt"] * a["charge"] / a["mass"]
        a["right_lim"] = 0.0
        a["w2pC"] = _W2PC
        a["dont_send"] = ["InjectorSource", "charge", "mass", "dens", "Immobile", "w2pC"]
        a["dont_keep"] = []
        self.flag_sorted = False

    def _init_data_on_dev(self):
        """particles.py:102-114 (empty SoA arrays of the step's precision)."""
        for arg in self._arg_names():
            self.DataDev[arg] = self.dev_arr(shape=(0,), dtype=self.comm.np_real)
