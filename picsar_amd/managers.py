"""The reference's zero-argument step managers on a Simulation's state.

PICSAR's drivers advance a step by calling managers that act on module state
(submain.F90:99-199):

    field_gathering_plus_particle_pusher   particle_pusher_manager_3d.F90:431
    push_laser_particles                   laser_pusher_manager_3d.F90:15
    particle_bcs                           particle_boundaries.F90:83
    pxr_particle_sorting                   sorting.F90:67
    pxrdepose_currents_on_grid_jxjyjz      current_deposition_manager_3d.F90:224
    current_bcs, efield_bcs, bfield_bcs    field_boundaries.F90:1861, 1757, 1808
    push_bfield, push_efield               maxwell_solver_manager.F90:44, 393
    pxrdepose_rho_on_grid, charge_bcs      charge_deposition_manager.F90:63,
                                           field_boundaries.F90:1910
    calc_diags                             diags.F90:63

``Managers(sim)`` exposes them under those names, acting on the fields and
tiled species that ``sim`` keeps on the GPU, so a driver written against the
managers runs unchanged, in the reference's order: the current is deposited
after the particle boundaries, from the wrapped positions.

The gather-and-push and the deposition managers run the tiled kernel of the
step in one phase each (FP64 or mixed-precision species); with periodic
particle faces ``Simulation.step`` runs both phases in one launch per
species.  The re-binning that ``particle_bcs`` performs also sorts every
tile, so ``pxr_particle_sorting`` has nothing left to do.
"""
from __future__ import annotations


class Managers:
    def __init__(self, sim):
        self.sim = sim

    def field_gathering_plus_particle_pusher(self):
        """Gather, push and position push of every species that is not frozen
        (particle_pusher_manager_3d.F90:431-672), tile by tile with the
        fields staged in LDS (FP64 or mixed-precision species)."""
        self.sim.gather_push()

    def push_laser_particles(self):
        for ant in self.sim.antennas:
            ant.push(self.sim)

    def particle_bcs(self):
        """Domain boundaries, tile re-binning and the per-tile sort."""
        self.sim.particle_bcs_and_sorting()

    def pxr_particle_sorting(self):
        """Done by particle_bcs: the re-binning writes every tile sorted."""

    def pxrdepose_currents_on_grid_jxjyjz(self):
        """J = 0, then the current of every species that deposits and of the
        antennas (current_deposition_manager_3d.F90:224-505), from the
        particles as particle_bcs left them, tile by tile with the current
        blocks accumulated in LDS."""
        self.sim.depose_currents()

    def current_bcs(self):
        self.sim.current_bcs()

    def push_bfield(self):
        self.sim.push_bfield()

    def bfield_bcs(self):
        self.sim.bfield_bcs()

    def push_efield(self):
        self.sim.push_efield()

    def efield_bcs(self):
        self.sim.efield_bcs()

    def pxrdepose_rho_on_grid(self):
        """rho = 0, then the charge density of every species that deposits
        (charge_deposition_manager.F90:63-250)."""
        self.sim.pxrdepose_rho_on_grid()

    def charge_bcs(self):
        """Periodic fold of the rho guards (field_boundaries.F90:1910-1930)."""
        self.sim.charge_bcs()

    def calc_diags(self):
        """calc_diags (diags.F90:63-115): the charge density and its guard
        fold (the reference's div E branch there is switched off)."""
        self.pxrdepose_rho_on_grid()
        self.charge_bcs()

    def calc_field_div(self):
        """div E into the simulation's ``dive`` array (diags.F90:129)."""
        return self.sim.calc_field_div()

    def get_norm_divErho(self) -> float:
        """||div E * eps0 - rho|| of the dive and rho arrays (diags.F90:1305)."""
        return self.sim.get_norm_divErho()

    def step(self, nst: int = 1):
        """submain.F90:99-199 with these managers (calc_diags included, as
        the reference calls it every step)."""
        for _ in range(int(nst)):
            self.field_gathering_plus_particle_pusher()
            self.push_laser_particles()
            self.particle_bcs()
            self.pxr_particle_sorting()
            self.pxrdepose_currents_on_grid_jxjyjz()
            self.current_bcs()
            self.push_bfield()
            self.bfield_bcs()
            self.push_efield()
            self.efield_bcs()
            self.push_bfield()
            self.bfield_bcs()
            self.calc_diags()
            self.sim.it += 1
        self.sim.check()
