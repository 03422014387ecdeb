"""Raw detector counts -> line integrals, corrected on the GPU at ingest.

The reference's loaders correct frames on the host and hand float32 line integrals
to ``SliceReconstructor.add_projections``:

* flexray.py:111-116 ``_read_binned`` (float32 block means, ``np.flipud``) and
  flexray.py:174-184 ``corrected``: ``-log(clip((raw - dark) * gain, 1e-3))`` with
  ``gain = 1 / max(flat - dark, 1)``, then the crop ``[r0:r1, c0:c1]`` (:201);
* hdf5.py:24-31 ``_bin`` and hdf5.py:77-78:
  ``-log(clip((raw - dark) / max(flat - dark, 1e-6), 1e-3))``.

``RawCorrection`` holds the binned dark / flat frames in those conventions and the
crop; ``SliceReconstructor.add_raw_projections(indices, frames, corr)`` uploads the
uint16 frames [k, H, W] (half the bytes of float32 line integrals) and
libminirecast_hip.so bins, flips, crops, corrects, filters and stores them.
"""
import numpy as np

STYLES = ("flexray", "hdf5")


class RawCorrection:
    """Flat/dark correction of raw frames.

    dark, flat: binned mean frames [Hb, Wb] as the loader builds them (flexray: from
    ``_read_binned``, so already flipped; hdf5: ``_bin(...).mean(0)``).
    frame_shape: (H, W) of the raw uint16 frames before binning.
    crop: (r0, r1, c0, c1) in binned coordinates (default: the whole binned frame).
    """

    def __init__(self, dark, flat, frame_shape, style="flexray", sample=1, crop=None, flip=None):
        if style not in STYLES:
            raise ValueError(f"unknown style {style!r}; expected one of {STYLES}")
        self.style = style
        self.sample = int(sample)
        if self.sample < 1:
            raise ValueError("sample must be >= 1")
        self.frame_shape = (int(frame_shape[0]), int(frame_shape[1]))
        hb, wb = self.frame_shape[0] // self.sample, self.frame_shape[1] // self.sample
        dark = np.asarray(dark, dtype=np.float32)
        flat = np.asarray(flat, dtype=np.float32)
        if dark.shape != (hb, wb) or flat.shape != (hb, wb):
            raise ValueError(f"dark/flat must be the binned frame shape {(hb, wb)}")
        self.flip = (style == "flexray") if flip is None else bool(flip)
        r0, r1, c0, c1 = crop if crop is not None else (0, hb, 0, wb)
        if not (0 <= r0 < r1 <= hb and 0 <= c0 < c1 <= wb):
            raise ValueError(f"crop {crop} outside the binned frame {(hb, wb)}")
        self.crop = (int(r0), int(r1), int(c0), int(c1))
        if style == "flexray":
            scale = 1.0 / np.maximum(flat - dark, 1.0)          # flexray.py:175 gain
            self.op = 0
        else:
            scale = np.maximum(flat - dark, 1e-6)               # hdf5.py:77 denominator
            self.op = 1
        self.dark = np.ascontiguousarray(dark[r0:r1, c0:c1])
        self.scale = np.ascontiguousarray(scale.astype(np.float32)[r0:r1, c0:c1])

    @property
    def shape(self):
        """(rows, cols) of the corrected, cropped detector."""
        r0, r1, c0, c1 = self.crop
        return r1 - r0, c1 - c0
