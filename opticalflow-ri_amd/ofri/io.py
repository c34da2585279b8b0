"""Frame reading and flow writing as the reference's example scripts do them.

Reference: examples/PyHSchunck_Fs3_4.py.  Frames are read as float32 grey
images (skimage.io.imread(fn).astype(np.float32), :129-130; Pillow here),
the HS smoothness alpha per pyramid level comes from the scripts' table of
bit depth x particle density (:63-122), and the accumulated flow is written
as a MATLAB .mat with the scripts' schema (save_flow, :35-51) so results
drop into the same PIV evaluation tooling.  Host-side only: nothing here
touches the GPU.
"""
import os
import re

import numpy as np

# HS alpha by (bit depth, particle density Ni) for the first level and the
# levels after it (PyHSchunck_Fs3_4.py:63-122; tuned on the parabolic,
# Rankine and uniform test flows with 3 px particles).
ALPHA_TABLE = {
    ("Bits08", "Ni01"): (6, 40), ("Bits08", "Ni06"): (21, 45),
    ("Bits08", "Ni12"): (30, 61), ("Bits08", "Ni16"): (34, 75),
    ("Bits10", "Ni01"): (20, 450), ("Bits10", "Ni06"): (77, 450),
    ("Bits10", "Ni12"): (119, 450), ("Bits10", "Ni16"): (131, 500),
    ("Bits12", "Ni01"): (82, 500), ("Bits12", "Ni06"): (325, 920),
    ("Bits12", "Ni12"): (470, 1260), ("Bits12", "Ni16"): (550, 1460),
}


def imread(path):
    """One frame as a float32 array: 8-bit ('L') and 16-bit ('I;16', the 10-
    and 12-bit sets) grey TIFFs keep their integer values.  Colour images are
    refused: the solvers take one intensity plane."""
    from PIL import Image
    with Image.open(path) as im:
        a = np.array(im)
    if a.ndim != 2:
        raise ValueError("%s: expected a single-channel image, got shape %s" % (path, a.shape))
    return a.astype(np.float32)


def read_sequence(paths):
    """Consecutive frame pairs (frame k, frame k+1) of an ordered list of
    files, each frame read once."""
    prev = None
    for p in paths:
        cur = imread(p)
        if prev is not None:
            if cur.shape != prev.shape:
                raise ValueError("%s: frame shape %s differs from %s" % (p, cur.shape, prev.shape))
            yield prev, cur
        prev = cur


def alpha_schedule(base_path, pyramidal_levels, k_levels):
    """The per-level HS alpha list the scripts pass to HSOpticalFlowAlgoAdapter:
    for each pyramid level l = 1..pyramidal_levels the table entry (first
    level or later) repeated k_levels times.  The bit depth and density are
    the BitsNN / NiNN components of base_path; an unlisted pair is an error
    (the scripts would fail on an undefined h)."""
    parts = re.split(r"[\\/]", os.path.normpath(base_path))
    bits = next((p for p in parts if re.fullmatch(r"Bits\d\d", p)), None)
    ni = next((p for p in parts if re.fullmatch(r"Ni\d\d", p)), None)
    if (bits, ni) not in ALPHA_TABLE:
        raise KeyError("no alpha table entry for %s / %s in %r" % (bits, ni, base_path))
    first, later = ALPHA_TABLE[(bits, ni)]
    out = []
    for level in range(1, int(pyramidal_levels) + 1):
        out += [first if level == 1 else later] * int(k_levels)
    return out


def save_flow(U, V, filename):
    """Write the flow as the scripts do (PyHSchunck_Fs3_4.py:35-51):
    velocities{u, v, iaWidth=1, iaHeight=1, margins{top,left,bottom,right}=0}
    and parameters{overlapFactor=1.0, imageHeight, imageWidth}."""
    import scipy.io
    U = np.asarray(U)
    V = np.asarray(V)
    margins = {"top": 0, "left": 0, "bottom": 0, "right": 0}
    results = {"u": U, "v": V, "iaWidth": 1, "iaHeight": 1, "margins": margins}
    parameters = {"overlapFactor": 1.0, "imageHeight": np.size(U, 0), "imageWidth": np.size(U, 1)}
    scipy.io.savemat(filename, mdict={"velocities": results, "parameters": parameters})


def load_flow(filename):
    """(U, V) back from a file written by save_flow."""
    import scipy.io
    m = scipy.io.loadmat(filename)
    vel = m["velocities"][0, 0]
    return vel["u"], vel["v"]
