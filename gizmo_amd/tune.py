"""Launch-shape autotuner: picks the library options that set kernel shapes for a workload by
timing short samples of it on the device.

The options (gz_set_option / gz_batch_set_option) change how the passes are launched, never
what they compute (every shape gives the same bits, which the parity tests check):

  few_rows      passes over at most this many rows run one workgroup per row (four wavefronts
                over the row's stencil or list), larger passes one lane per row or one
                wavefront per cell; the crossover depends on the rows' list lengths and on how
                many rows fill the chip
  row_schedule  the force pass's lanes take each block of this many rows longest list first
                (0: the order of the rows); pays when lists differ in length within a wave
  speculate     the run loop launches the passes behind the first density round before its
                counters are read (one host wait fewer per step, a repeat of the passes when
                rows still iterate)

``tune(sample, space)`` times ``sample(options)`` (seconds, one call per candidate and repeat,
best of ``reps``) over the candidates of each option in turn, keeping the best value of an
option before the next one is tuned (coordinate descent from the defaults, which are the
first value of each list).  The samples are the caller's: a few steps of a fresh job
(``job_sample``), a few steps of a fresh batch (``batch_sample``), or full passes over a
resident state (``pass_sample``).
"""
import time

SINGLE_BOX_SPACE = {
    "few_rows": [8192, 2048, 32768, 0],
    "row_schedule": [2048, 0, 512, 128],
    "speculate": [1, 0],
}
BATCH_SPACE = {"few_rows": [8192, 2048, 32768, 0]}


def tune(sample, space, reps=2, log=None):
    """Coordinate descent over `space` ({option: [default, other values...]}); returns
    (best options, table of (options, seconds) for every timed candidate)."""
    best = {k: v[0] for k, v in space.items()}
    table = []
    best_t = None
    for k, values in space.items():
        for v in values:
            opts = dict(best, **{k: v})
            if best_t is not None and opts == best:
                continue  # timed already
            t = min(sample(opts) for _ in range(reps))
            table.append((dict(opts), t))
            if log:
                log(f"tune: {opts} {1e3 * t:.3f} ms")
            if best_t is None or t < best_t:
                best, best_t = opts, t
    return best, table


def job_sample(make_hydro, integ, steps):
    """A sample of the run loop: a fresh context from make_hydro(), the options set,
    gz_run_begin untimed, then `steps` sync points timed."""
    def run(opts):
        g = make_hydro()
        for k, v in opts.items():
            g.set_option(k, v)
        g.run_begin(integ)
        g.synchronize()
        t0 = time.perf_counter()
        g.run_steps(steps)
        g.synchronize()
        dt = time.perf_counter() - t0
        g.close()
        return dt
    return run


def pass_sample(g, reps=1):
    """A sample of full hot-path passes over the resident state of context g (the radii
    converged already, so every pass computes the same numbers)."""
    def run(opts):
        for k, v in opts.items():
            g.set_option(k, v)
        g.hydro_pass()
        g.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            g.hydro_pass()
        g.synchronize()
        return (time.perf_counter() - t0) / reps
    return run


def batch_sample(make_batch, steps):
    """A sample of a batch: a fresh batch from make_batch() (boxes added, not begun), the
    options set, begin untimed, then `steps` shared steps timed."""
    def run(opts):
        b = make_batch()
        for k, v in opts.items():
            b.set_option(k, v)
        b.begin()
        b.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.step()
        b.synchronize()
        dt = time.perf_counter() - t0
        b.close()
        return dt
    return run
