"""Board-power sampling around a command or a region of a program.

The counterpart of the reference's power scripts (scripts/power/power_520N.sh,
scripts/power/power_PACD5005.sh: start the benchmark in the background, read
the board's total power in a loop until it exits, one reading per line of
powermeasure.csv), for AMD Instinct GPUs through the AMD SMI library.

    python -m convfpga_amd.power [-o powermeasure.csv] [-i 0.01] [-g BDF] -- <command> [args...]

runs <command> as a child process, samples the GPU's socket power every
interval until the child exits, writes `time_s,power_W` lines, prints the
mean and peak power and the energy the GPU's own energy counter accumulated
over the run, and exits with the child's status.  `PowerMonitor` does the
same around a region of a Python program (bench.py --power).
"""
import argparse
import csv
import glob
import os
import subprocess
import sys
import threading
import time


class PowerUnavailable(RuntimeError):
    pass


class SmiReader:
    """Socket power (W) and the energy counter (J) of one GPU through amdsmi."""

    def __init__(self, bdf=None):
        try:
            import amdsmi
        except ImportError as e:  # pragma: no cover - the image ships amdsmi
            raise PowerUnavailable("amdsmi is not importable: %s" % e)
        self._smi = amdsmi
        try:
            amdsmi.amdsmi_init()
        except Exception as e:
            raise PowerUnavailable("amdsmi_init failed: %s" % e)
        try:
            handles = amdsmi.amdsmi_get_processor_handles()
        except Exception as e:
            amdsmi.amdsmi_shut_down()
            raise PowerUnavailable("no GPU handles: %s" % e)
        if not handles:
            amdsmi.amdsmi_shut_down()
            raise PowerUnavailable("amdsmi sees no GPU")
        self.seen = [amdsmi.amdsmi_get_gpu_device_bdf(h).lower() for h in handles]
        if not bdf:  # the GPU this process may open, when there is exactly one
            mine = visible_gpu_bdfs()
            bdf = mine[0] if len(mine) == 1 else None
        self.handle, self.bdf = handles[0], self.seen[0]
        if bdf:
            for h, b in zip(handles, self.seen):
                if b.startswith(bdf.lower()):  # amdsmi prints domain:bus:device.function
                    self.handle, self.bdf = h, b
                    break
            else:
                amdsmi.amdsmi_shut_down()
                raise PowerUnavailable("no GPU with bus id %s among %s" % (bdf, ", ".join(self.seen)))
        elif len(handles) > 1:
            amdsmi.amdsmi_shut_down()
            raise PowerUnavailable("%d GPUs (%s): name one with -g" % (len(handles), ", ".join(self.seen)))

    def power_w(self):
        info = self._smi.amdsmi_get_power_info(self.handle)
        for key in ("current_socket_power", "average_socket_power", "socket_power"):
            v = info.get(key, "N/A")
            if v != "N/A" and v:
                return float(v)
        return None

    def energy_j(self):
        try:
            e = self._smi.amdsmi_get_energy_count(self.handle)
        except Exception:
            return None
        return e["energy_accumulator"] * e["counter_resolution"] * 1e-6  # counter unit: microjoules

    def close(self):
        try:
            self._smi.amdsmi_shut_down()
        except Exception:
            pass


def visible_gpu_bdfs():
    """PCI bus ids of the GPUs whose render nodes this process can open."""
    out = []
    for node in sorted(glob.glob("/dev/dri/renderD*")):
        if os.access(node, os.R_OK | os.W_OK):
            dev = os.path.realpath("/sys/class/drm/%s/device" % os.path.basename(node))
            if os.path.basename(dev).count(":") == 2:
                out.append(os.path.basename(dev).lower())
    return out


def torch_device_bdf(index=0):
    """PCI bus id of torch's cuda:index, as amdsmi prints it (domain:bus:dev.fn)."""
    try:
        import torch
        p = torch.cuda.get_device_properties(index)
        return "%04x:%02x:%02x" % (p.pci_domain_id, p.pci_bus_id, p.pci_device_id)
    except Exception:
        return None


class PowerMonitor:
    """Samples power on a background thread between start() and stop();
    `reader` is any object with power_w() and energy_j() (SmiReader on the GPU)."""

    def __init__(self, reader, interval_s=0.01):
        self.reader, self.interval = reader, interval_s
        self.samples = []  # (seconds since start, watts)
        self._stop = threading.Event()
        self._thread = None

    def _run(self):
        while not self._stop.is_set():
            w = self.reader.power_w()
            if w is not None:
                self.samples.append((time.perf_counter() - self.t0, w))
            self._stop.wait(self.interval)

    def start(self):
        self.t0 = time.perf_counter()
        self.e0 = self.reader.energy_j()
        self._thread = threading.Thread(target=self._run, daemon=True)
        self._thread.start()
        return self

    def stop(self):
        self._stop.set()
        self._thread.join()
        self.t1 = time.perf_counter()
        self.e1 = self.reader.energy_j()
        return self.summary()

    def summary(self):
        ws = [w for _, w in self.samples]
        dt = self.t1 - self.t0
        out = {"seconds": dt, "samples": len(ws),
               "mean_W": sum(ws) / len(ws) if ws else None, "max_W": max(ws) if ws else None}
        if self.e0 is not None and self.e1 is not None and self.e1 >= self.e0 and dt > 0:
            out["energy_J"] = self.e1 - self.e0
            out["energy_mean_W"] = (self.e1 - self.e0) / dt
        return out

    def write_csv(self, path):
        with open(path, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(["time_s", "power_W"])
            for t, p in self.samples:
                w.writerow(["%.4f" % t, "%.1f" % p])


def main(argv=None, reader=None):
    ap = argparse.ArgumentParser(prog="python -m convfpga_amd.power",
                                 description="sample GPU power while a command runs")
    ap.add_argument("-o", "--output", default="powermeasure.csv", help="CSV of time_s,power_W (default %(default)s)")
    ap.add_argument("-i", "--interval", type=float, default=0.01, help="seconds between samples")
    ap.add_argument("-g", "--gpu", default=None,
                    help="PCI bus id of the GPU (default: the one GPU this process can open)")
    ap.add_argument("command", nargs=argparse.REMAINDER, help="-- command [args...]")
    a = ap.parse_args(argv)
    cmd = a.command[1:] if a.command[:1] == ["--"] else a.command
    if not cmd:
        ap.error("no command given")
    try:
        rd = reader or SmiReader(a.gpu)
    except PowerUnavailable as e:
        print("power: %s" % e, file=sys.stderr)
        return 2
    if reader is None:
        print("power: GPU %s (amdsmi sees %d)" % (rd.bdf, len(rd.seen)), file=sys.stderr)
    mon = PowerMonitor(rd, a.interval).start()
    try:
        rc = subprocess.call(cmd)
    finally:
        s = mon.stop()
        mon.write_csv(a.output)
        if reader is None:
            rd.close()
    msg = "power: %d samples over %.2f s" % (s["samples"], s["seconds"])
    if s["mean_W"] is not None:
        msg += ", mean %.1f W, peak %.1f W" % (s["mean_W"], s["max_W"])
    if "energy_J" in s:
        msg += ", %.1f J by the energy counter (%.1f W mean)" % (s["energy_J"], s["energy_mean_W"])
    print(msg + " -> " + a.output, file=sys.stderr)
    return rc


if __name__ == "__main__":
    sys.exit(main())
