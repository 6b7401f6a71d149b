"""Single-GPU stand-in for the reference's ParallelMPI object.

The reference threads a ``ParallelMPI`` instance through every operator for
logging (``root_print``), fatal errors (``kill`` -> MPI_Abort) and
reductions (``ParallelMPI.pyx:17-241``).  One GPU owns the whole domain here,
so the reductions are device kernels and this object only keeps the logging
and error surface: ``kill`` raises instead of aborting the job.
"""


class ParallelMPI:
    rank = 0
    size = 1

    def __init__(self, namelist=None, verbose=False):
        self.verbose = verbose

    def root_print(self, msg):
        if self.verbose:
            print(msg, flush=True)

    def kill(self, msg="killed"):
        raise RuntimeError(msg)

    def barrier(self):
        return None


class NetCDFIO_Stats:
    """Statistics output is outside the device path; calls are accepted and ignored."""

    def add_profile(self, *a, **k):
        pass

    def add_ts(self, *a, **k):
        pass

    def add_reference_profile(self, *a, **k):
        pass

    def write_profile(self, *a, **k):
        pass

    def write_ts(self, *a, **k):
        pass
