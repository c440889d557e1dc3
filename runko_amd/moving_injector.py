"""Moving particle injector of the shock setup (runko/moving_injector.py:4-54).

The injection front starts at injloc and advances with beta_inj * cfl per
lap.  Every n_inj laps (never on lap 0) the cells of the x stripe between the
previous front, pulled back by the upstream drift beta_flow * n_inj * cfl and
never behind the wall, and the new front receive ppc rounds of every species'
generator through Tile.batch_inject_in_x_stripe.  The front stops moving once
it comes within Lx_margin of Lx.
"""


class MovingInjector:
    def __init__(self, injloc, beta_inj, beta_flow, cfl, n_inj=1, walloc=0.0, Lx_margin=10.0, Lx=None):
        self.injloc = injloc
        self.beta_inj = beta_inj
        self.beta_flow = beta_flow
        self.cfl = cfl
        self.n_inj = n_inj
        self.walloc = walloc
        self.Lx_margin = Lx_margin
        self.Lx = Lx
        self.moving = True

    def inject(self, simulation, pgens, ppc):
        """Call once per lap; pgens is a list of (species, generator) pairs."""
        if not self.moving or simulation.lap == 0 or simulation.lap % self.n_inj != 0:
            return
        span = self.n_inj * self.cfl
        left = max(self.injloc - self.beta_flow * span, self.walloc)
        right = self.injloc + self.beta_inj * span
        if self.Lx is not None and right >= self.Lx - self.Lx_margin:
            self.moving = False
            return
        for tile in simulation.local_tiles():
            for _ in range(ppc):
                for species, pgen in pgens:
                    tile.batch_inject_in_x_stripe(species, pgen, left, right)
        self.injloc = right
