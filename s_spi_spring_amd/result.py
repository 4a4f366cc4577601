"""Result of a memetic reconstruction, with the reference's Result surface
(spring/result.py:44-790) for everything that does not plot: best and average
density and support, statistics, settings, pattern, cropping to the support,
residuals, the reconstructed pattern, the phase-retrieval transfer function,
alignment of another density to this one, and the status print.

save()/load() write and read the reference's HDF5 layout
(spring/result.py:115-259) when h5py is installed; h5py is not part of this
image, so that layout is not exercised by the tests here.
"""
from __future__ import annotations

import os
from copy import deepcopy

import numpy as np


class Result:
    def __init__(self, pid: list = None, workdir: str = "./", tag: str = ""):
        self.pattern = None
        self.density_best = None
        self.support_best = None
        self.density_avg = None
        self.support_avg = None
        self.pid = None
        self.stats = None
        self.settings = None
        self.metadata = None
        self.density_all = None
        self.support_all = None
        self.proc_id = os.getpid()
        self.step = None
        self.total = None
        self.fixsize = 0
        if pid is not None:
            self.load(pid, workdir, tag)
        self.x_min, self.y_min = None, None
        self.x_max, self.y_max = None, None
        self.margsize = 0.05

    def set(self, pattern, density_best, support_best, density_avg, support_avg, pid, stats, settings, metadata=None,
            density_all=None, support_all=None, step=None, total=None):
        self.pattern = pattern
        self.density_best = density_best
        self.support_best = support_best
        self.density_avg = density_avg
        self.support_avg = support_avg
        self.pid = pid
        self.stats = stats
        self.settings = settings
        self.metadata = {} if metadata is None else metadata
        self.density_all = density_all
        self.support_all = support_all
        self.step = step
        self.total = total
        self.x_min = None

    def copy(self, inres):
        for k in ("pattern", "density_best", "support_best", "density_avg", "support_avg", "pid", "stats", "settings",
                  "metadata", "density_all", "support_all"):
            setattr(self, k, deepcopy(getattr(inres, k)))
        self.step, self.total, self.x_min = inres.step, inres.total, None

    # --- files (spring/result.py:115-259) -------------------------------------
    def get_filename(self, pid: list, tag=""):
        out = "".join("{:d}_".format(int(p)) for p in pid)
        if len(tag) > 0 and tag != "*":
            out += tag + "_"
        if tag == "*":
            out += "step{:04d}_".format(self.step)
        return out + "rec.h5"

    def save(self, workdir: str = "./", tag: str = ""):
        fullname = os.path.join(workdir, self.get_filename(self.pid, tag))
        try:
            self.savefile(fullname)
        except Exception:
            print("Failed to save file {}".format(fullname))

    def savefile(self, fname: str):
        import h5py
        with h5py.File(fname, "w") as f:
            f.create_dataset("pid", data=self.pid)
            f.create_dataset("settings", data=self.settings.getjson())
            for k in ("pattern", "density_best", "support_best", "density_avg", "support_avg"):
                f.create_dataset(k, data=np.nan_to_num(getattr(self, k)))
            if self.density_all is not None:
                f.create_dataset("density_all", data=self.density_all)
            if self.support_all is not None:
                f.create_dataset("support_all", data=self.support_all)
            for key in self.metadata:
                f.create_dataset("metadata/" + str(key), data=self.metadata[key])
            if self.stats is not None:
                sd = self.stats.get()
                for k1 in sd:
                    for k2 in sd[k1]:
                        f.create_dataset("stats/{}/{}".format(k1, k2), data=sd[k1][k2])
            f.create_dataset("process/generation", data=self.step)
            f.create_dataset("process/totgenerations", data=self.total)
            f.create_dataset("process/proc_id", data=self.proc_id)

    def load(self, pid: list, workdir: str = "./", tag: str = ""):
        fullname = os.path.join(workdir, self.get_filename(pid, tag))
        try:
            self.loadfile(fullname)
        except Exception:
            print("Failed to load file {}".format(fullname))

    def loadfile(self, fname: str):
        import h5py

        from .settings import Settings
        from .stats import Stats
        with h5py.File(fname, "r") as f:
            for k in ("pattern", "density_best", "support_best", "density_avg", "support_avg", "pid"):
                setattr(self, k, f[k][()])
            self.settings = Settings()
            self.settings.setjson(jsonstr=str(f["settings"][()].decode("UTF-8")))
            self.metadata = {k: f["metadata"][k][()] for k in f["metadata"].keys()} if "metadata" in f.keys() else {}
            sd = {k1: {k2: f["stats"][k1][k2][()] for k2 in f["stats"][k1].keys()} for k1 in f["stats"].keys()}
            self.stats = Stats()
            self.stats.set(sd)
            if "density_all" in f.keys():
                self.density_all = f["density_all"][()]
            if "support_all" in f.keys():
                self.support_all = f["support_all"][()]
            self.step = f["process/generation"][()]
            self.total = f["process/totgenerations"][()]
            self.proc_id = f["process/proc_id"][()]
        self.x_min, self.y_min = None, None
        self.x_max, self.y_max = None, None

    # --- access (spring/result.py:263-408) -------------------------------------
    def print_status(self):
        print("Generation {:d}/{:d}".format(self.step, self.total))
        self.stats.print()

    def get_crop(self):
        """Square bounding box of the best support (padded to the larger side,
        or `fixsize` wide), with margins of `margsize` of its width."""
        coords = np.argwhere(self.support_best)
        self.x_min, self.y_min = coords.min(axis=0)
        self.x_max, self.y_max = coords.max(axis=0)
        if self.fixsize > 0:
            midx, midy = (self.x_min + self.x_max) // 2, (self.y_min + self.y_max) // 2
            self.x_min = midx - self.fixsize // 2
            self.x_max = self.x_min + self.fixsize
            self.y_min = midy - self.fixsize // 2
            self.y_max = self.y_min + self.fixsize
        elif self.x_max - self.x_min < self.y_max - self.y_min:
            c = (self.x_max + self.x_min) // 2
            self.x_min = c - (self.y_max - self.y_min) // 2
            self.x_max = c + (self.y_max - self.y_min) // 2
        elif self.x_max - self.x_min > self.y_max - self.y_min:
            c = (self.y_max + self.y_min) // 2
            self.y_min = c - (self.x_max - self.x_min) // 2
            self.y_max = c + (self.x_max - self.x_min) // 2
        if self.margsize >= 0:
            self.margins = int((self.x_max - self.x_min) * self.margsize)

    def crop(self, data):
        if self.x_min is None:
            self.get_crop()
        return np.pad(data[self.x_min:self.x_max + 1, self.y_min:self.y_max + 1], self.margins)

    def get_density_pop(self, index, crop: bool = True):
        if self.density_all is not None and 0 <= index < len(self.density_all):
            return self.crop(self.density_all[index]) if crop else np.copy(self.density_all[index])
        return None

    def get_density(self, crop: bool = True, indens=None):
        out = self.density_avg if indens is None else indens
        return self.crop(out) if crop else np.copy(out)

    def get_support(self, crop: bool = True):
        return self.crop(self.support_avg) if crop else np.copy(self.support_avg)

    def get_density_best(self, crop: bool = True):
        return self.crop(self.density_best) if crop else np.copy(self.density_best)

    def get_support_best(self, crop: bool = True):
        return self.crop(self.support_best) if crop else np.copy(self.support_best)

    def get_stats(self):
        return self.stats.get()

    def get_settings(self):
        return self.settings

    def get_pattern(self):
        return self.pattern

    def get_rec_pattern(self, indens=None):
        """|FFT(conj rho)|^2, centred: the intensity of the reconstruction in the
        pattern's orientation (spring/result.py:402-408)."""
        d = self.get_density(crop=False, indens=indens)
        return np.abs(np.fft.fftshift(np.fft.fft2(np.conj(d)))) ** 2

    def get_residuals(self, indens=None):
        return np.abs(self.pattern - self.get_rec_pattern(indens))

    def get_prtf(self):
        """Phase retrieval transfer function of the average density
        (spring/result.py:688-716): radial profile of min(|FFT(conj rho)| / sqrt(I), 1)
        over the measured pixels; q normalised to [0, 1)."""
        from . import spring_core
        rec = np.abs(np.fft.fftshift(np.fft.fft2(np.conj(self.get_density(crop=False)))))
        orig = np.copy(self.pattern).astype(np.float64)
        orig[orig < 0] = 0
        orig = orig ** 0.5
        ratio = -1.0 * np.ones(orig.shape)
        ratio[orig > 0] = rec[orig > 0] / orig[orig > 0]
        ratio[ratio > 1] = 1
        center = [self.pattern.shape[0] / 2, self.pattern.shape[1] / 2]
        nangles = self.pattern.shape[0] // 2
        dtheta = 1 / nangles
        angles = np.arange(0, 1, dtheta)
        prof = np.array(spring_core.pattern.get_radial_profile(ratio, dtheta, center, list(angles), [0, 360]))
        return angles[prof > 0], prof[prof > 0]

    def reshift(self, density, crop=True):
        """`density` aligned to the average density (spring/result.py:748-790):
        of the density and its twin (point reflection, conjugated), the one with
        the larger peak of the FFT cross-correlation (the twin on a tie), moved
        cyclically by the peak's offset and rotated by the peak's phase."""
        ref_ft = np.fft.fft2(self.get_density(crop=False))
        d = np.asarray(density)
        best = None
        for cand in (d, np.conj(d[::-1, ::-1])):
            corr = np.fft.ifft2(np.conj(ref_ft) * np.fft.fft2(cand))
            k = np.unravel_index(np.argmax(np.abs(corr)), corr.shape)
            if best is None or abs(corr[k]) >= abs(best[1]):
                best = (cand, corr[k], k)
        cand, peak, k = best
        out = np.roll(cand, (-k[0], -k[1]), axis=(0, 1)) * np.exp(-1j * np.angle(peak))
        return self.crop(out) if crop else out
