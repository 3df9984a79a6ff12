"""The reference's running statistics (stats.F90:152-271): used by the MD
display and the post-processor."""
import math

STATS_BLOCKSIZE = 10000    # defaults.F90:138


class Stat:
    """stats_update / stats_getfield (stats.F90:152-271) without umbrella
    weights: instantaneous value, block average (restarts after each
    block of STATS_BLOCKSIZE updates), cumulative average, and
    std = sqrt(<x^2> - <x>^2) from running averages."""

    def __init__(self, blocksize=STATS_BLOCKSIZE):
        self.blocksize = blocksize
        self.blockiters = 0
        self.cumiters = 0
        self.inst = self.blockavg = self.cumavg = self.cumsqavg = self.cumdevsq = 0.0

    def update(self, value):
        self.inst = value
        self.blockiters += 1
        self.blockavg = (self.blockavg * (self.blockiters - 1.0) + value) / (self.blockiters * 1.0)
        if self.blockiters == self.blocksize:
            self.blockiters = 0
        self.cumiters += 1
        self.cumavg = (self.cumavg * (self.cumiters - 1.0) + value) / (self.cumiters * 1.0)
        self.cumsqavg = (self.cumsqavg * (self.cumiters - 1.0) + value * value) / (self.cumiters * 1.0)
        self.cumdevsq = self.cumsqavg - self.cumavg ** 2

    def std(self):
        return 0.0 if self.cumdevsq < 0.0 else math.sqrt(self.cumdevsq)

    def row(self):
        return (self.inst, self.cumavg, self.blockavg, self.std())
