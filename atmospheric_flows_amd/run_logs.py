"""The reference driver's run logs for a run driven through EulerOperator.

EulerSolver writes, into its output directory (Atmospheric_Flow.cc:246-255):
  - n_dofs_velocity.dat, n_dofs_temperature.dat, n_dofs_density.dat: the DoF count of each field,
    one line each, at set-up (:318-322);
  - error_analysis_vel.dat, error_analysis_rho.dat, error_analysis_pres.dat: after every analysis,
    the L2 error against the exact field and the error relative to the exact field's L2 norm,
    one line each (:925-985);
  - time_analysis_<n>proc.dat: deal.II's TimerOutput summary (cpu_and_wall_times) of the run's
    sections (:249, 281, 418, 432, 546, 682, 790, 887, 927).
RunLogs writes the same files. Host sections (set-up, steps, output, analysis) are timed on the host,
with the device synchronised at their ends; the four phases inside afl_imex_step (Update density,
Fixed point pressure, Update velocity, Update pressure) come from the library's HIP-event section
timer, so their rows hold device wall time and appear in the wall-time table only. The L2 norms use
the operators' 2k+1 Gauss points per direction (the reference integrates with k+1).
"""
import contextlib
import os
import time

__all__ = ["RunLogs"]


class RunLogs:
    ANALYSIS = "Analysis results: computing errrors"  # the reference's section name, spelling included

    def __init__(self, op, directory, nproc=1):
        self.op = op
        self.dir = directory
        self.nproc = nproc
        os.makedirs(directory, exist_ok=True)
        n = op.N
        for name, count in (("velocity", op.dim * n), ("temperature", n), ("density", n)):
            with open(os.path.join(directory, f"n_dofs_{name}.dat"), "w") as f:
                f.write(f"{count}\n")
        for name in ("vel", "rho", "pres"):
            open(os.path.join(directory, f"error_analysis_{name}.dat"), "w").close()
        self.sections = {}  # host section -> [calls, cpu seconds, wall seconds]
        self.t0, self.c0 = time.perf_counter(), time.process_time()
        op.section_timer_start()

    @contextlib.contextmanager
    def section(self, name):
        """Times a host section (TimerOutput::Scope); the device is synchronised at both ends."""
        torch = self.op.torch
        torch.cuda.synchronize()
        t, c = time.perf_counter(), time.process_time()
        try:
            yield
        finally:
            torch.cuda.synchronize()
            rec = self.sections.setdefault(name, [0, 0.0, 0.0])
            rec[0] += 1
            rec[1] += time.process_time() - c
            rec[2] += time.perf_counter() - t

    def analyze(self, rho, u, p, rho_exact, u_exact, p_exact):
        """Appends (L2 error, relative L2 error) of each field to its error_analysis file."""
        out = {}
        with self.section(self.ANALYSIS):
            for name, ncomp, a, b in (("vel", self.op.dim, u, u_exact), ("rho", 1, rho, rho_exact),
                                      ("pres", 1, p, p_exact)):
                err, ref = self.op.l2_difference(ncomp, a, b)
                out[name] = (err, err / ref)
        for name, (err, relerr) in out.items():
            with open(os.path.join(self.dir, f"error_analysis_{name}.dat"), "a") as f:
                f.write(f"{err!r}\n{relerr!r}\n")
        return out

    def write_time_analysis(self):
        """Writes time_analysis_<n>proc.dat: the CPU-time table of the host sections and the wall-time table
        of the host sections and the device phases of afl_imex_step."""
        self.op.torch.cuda.synchronize()
        cpu_total = time.process_time() - self.c0
        wall_total = time.perf_counter() - self.t0
        device = []
        for i, name in enumerate(self.op.SECTIONS):
            calls, ms = self.op.section_timer_result(i)
            if calls:
                device.append((name, calls, ms * 1e-3))
        cpu_rows = [(n, r[0], r[1]) for n, r in self.sections.items()]
        wall_rows = [(n, r[0], r[2]) for n, r in self.sections.items()] + device
        text = _table("Total CPU time elapsed since start", "CPU time", cpu_total, cpu_rows)
        text += "\n" + _table("Total wallclock time elapsed since start", "wall time", wall_total, wall_rows)
        path = os.path.join(self.dir, f"time_analysis_{self.nproc}proc.dat")
        with open(path, "w") as f:
            f.write(text)
        return path


def _table(title, column, total, rows):
    """One TimerOutput summary table (deal.II's layout)."""
    sep = "+---------------------------------------------+------------+------------+\n"
    mid = "+---------------------------------+-----------+------------+------------+\n"
    out = "\n" + sep
    out += f"| {title:<44s}|{total:>10.3g}s |            |\n"
    out += "|                                             |            |            |\n"
    out += f"| Section                         | no. calls | {column:>10s} | % of total |\n"
    out += mid
    for name, calls, secs in sorted(rows):
        pct = 100.0 * secs / total if total > 0 else 0.0
        out += f"| {name[:32]:<32s}|{calls:>10d} |{secs:>10.3g}s |{pct:>10.2g}% |\n"
    out += mid
    return out
