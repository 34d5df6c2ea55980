"""Saving and loading a DFN with its solution (the checkpoint of ``andfn/dfn.py:404-609``).

The saved content follows AnDFN's ``save_dfn``: every field of the consolidated element and
fracture records under ``elements/properties/<field>`` and ``fractures/properties/<field>``,
and the index arrays (``label``, ``_id``, ``_type`` of the elements; ``label``, ``_id`` of the
fractures) under ``elements/index`` and ``fractures/index``.  A loaded DFN carries the solved
coefficients, coefficient counts, discharges, errors and fracture constants, so its next solve
starts from them (``consolidate_dfn`` copies ``coef`` and ``q`` into the records, as
``andfn/dfn.py:727,734`` does): a warm start.

Two containers:

* ``.h5`` (needs h5py): AnDFN's own layout, one dataset per record and field
  (``elements/properties/<field>/<i>``), readable by AnDFN's ``load_dfn``, and AnDFN's files
  readable here;
* ``.npz``: the same datasets stacked per field (``elements/properties/<field>`` holds the
  field of every record in id order), written with numpy and loaded without pickling.

Beyond AnDFN's layout the files hold ``meta/*`` (the solver constants, ``discharge_int``, the
record capacities), and the loader restores each fracture's element order from the saved
``elements`` lists (AnDFN's loader appends elements in id order), so a reloaded network
reproduces the saved records exactly.
"""

import os

import numpy as np

from .dtypes import (
    BOUNDING_CIRCLE,
    CONST_HEAD_LINE,
    IMP_CIRCLE,
    IMP_LINE,
    INTERSECTION,
    WELL,
    dtype_constants,
)


def _filename(filename):
    ext = os.path.splitext(filename)[1]
    if ext == "":
        return filename + ".h5", ".h5"
    if ext not in (".h5", ".npz"):
        raise ValueError(f"Unsupported file extension: {ext}. Supported extensions are .h5 and .npz.")
    return filename, ext


def _h5py():
    try:
        import h5py
    except ImportError as e:  # pragma: no cover - h5py is optional
        raise ImportError("saving or loading .h5 files needs h5py; use a .npz file instead") from e
    return h5py


def _content(dfn):
    """Arrays to save: (group, field) -> stacked array over the records, in id order."""
    if dfn.elements_struc_array_hpc is None or dfn.fractures_struc_array_hpc is None:
        dfn.consolidate_dfn()
    ea, fa = dfn.elements_struc_array_hpc, dfn.fractures_struc_array_hpc
    out = {}
    for name in ea.dtype.names:
        out[("elements/properties", name)] = ea[name]
    for name in fa.dtype.names:
        out[("fractures/properties", name)] = fa[name]
    ap = np.array([np.nan if f.aperture is None else float(f.aperture) for f in dfn.fractures])
    out[("fractures/properties", "aperture")] = ap
    out[("elements/index", "label")] = np.array([str(e.label) for e in dfn.elements]).astype("S")
    out[("elements/index", "_id")] = ea["_id"].copy()
    out[("elements/index", "_type")] = ea["_type"].copy()
    out[("fractures/index", "label")] = np.array([str(f.label) for f in dfn.fractures]).astype("S")
    out[("fractures/index", "_id")] = fa["_id"].copy()
    meta = {"discharge_int": np.int64(dfn.discharge_int), "max_ncoef": np.int64(dfn.max_ncoef),
            "max_elements": np.int64(dfn.max_elements), "label": np.bytes_(str(dfn.label))}
    for name in dtype_constants.names:
        meta["constants_" + name] = np.asarray(dfn.constants[name]).reshape(())
    for k, v in meta.items():
        out[("meta", k)] = np.asarray(v)
    return out


def save_dfn(dfn, filename, overwrite=False):
    """Write the DFN and its solution to ``filename`` (``andfn/dfn.py:404-458``).  Returns the
    file name written, or None when the file exists and ``overwrite`` is False (AnDFN logs a
    warning and returns)."""
    import logging

    filename, ext = _filename(filename)
    if os.path.exists(filename) and not overwrite:
        logging.getLogger("andfn").warning(
            f"The file {filename} already exists. To overwrite the existing file set the argument 'overwrite' to True.")
        return None
    content = _content(dfn)
    if ext == ".npz":
        np.savez(filename, **{f"{g}/{k}": v for (g, k), v in content.items()})
    else:
        h5py = _h5py()
        with h5py.File(filename, "w") as hf:
            for (g, k), v in content.items():
                if g.endswith("/properties"):
                    grp = hf.require_group(f"{g}/{k}")
                    for i, row in enumerate(v):
                        grp.create_dataset(f"{i}", data=row)
                else:
                    hf.create_dataset(f"{g}/{k}", data=v)
    logging.getLogger("andfn").info(f"Saved DFN to {filename}")
    return filename


class _Reader:
    """Field access over either container: ``get(group, field, i)`` and whole-field arrays."""

    def __init__(self, filename, ext):
        self.ext = ext
        if ext == ".npz":
            self.z = np.load(filename, allow_pickle=False)
            self.keys = set(self.z.files)
        else:
            self.z = _h5py().File(filename, "r")
            self.keys = None

    def has(self, group, field, i=None):
        if self.ext == ".npz":
            return f"{group}/{field}" in self.keys
        key = f"{group}/{field}" if i is None else f"{group}/{field}/{i}"
        return key in self.z

    def index(self, group, field):
        if self.ext == ".npz":
            return self.z[f"{group}/{field}"]
        return self.z[f"{group}/{field}"][()]

    def get(self, group, field, i):
        if self.ext == ".npz":
            return self.z[f"{group}/{field}"][i]
        return self.z[f"{group}/{field}/{i}"][()]

    def close(self):
        if self.ext != ".npz":
            self.z.close()


def load_dfn(dfn, filename):
    """Read fractures and elements with their solution into ``dfn`` (``andfn/dfn.py:460-609``)."""
    import logging

    from .dfn import (
        BoundingCircle,
        ConstantHeadLine,
        Fracture,
        ImpermeableCircle,
        ImpermeableLine,
        Intersection,
        Well,
    )

    filename, ext = _filename(filename)
    if not os.path.exists(filename):
        raise FileNotFoundError(f"The file {filename} does not exist.")
    logging.getLogger("andfn").info(f"Loading DFN from {filename}")
    rd = _Reader(filename, ext)
    try:
        if rd.has("meta", "discharge_int"):
            dfn.discharge_int = int(rd.index("meta", "discharge_int"))
            for name in dtype_constants.names:
                if rd.has("meta", "constants_" + name):
                    dfn.constants[name] = rd.index("meta", "constants_" + name)
        FP, EP = "fractures/properties", "elements/properties"
        flabels = rd.index("fractures/index", "label")
        fracs = []
        for i in range(len(flabels)):
            ap = float(rd.get(FP, "aperture", i)) if rd.has(FP, "aperture", i) else 1e-6
            basis = (np.asarray(rd.get(FP, "normal", i), dtype=np.float64),
                     np.asarray(rd.get(FP, "x_vector", i), dtype=np.float64),
                     np.asarray(rd.get(FP, "y_vector", i), dtype=np.float64))
            f = Fracture(flabels[i].decode(), float(rd.get(FP, "t", i)), float(rd.get(FP, "radius", i)),
                         np.asarray(rd.get(FP, "center", i), dtype=np.float64), basis[0],
                         aperture=None if ap != ap else ap, elements=False, basis=basis)
            f.constant = float(rd.get(FP, "constant", i))
            fracs.append(f)
        elabels = rd.index("elements/index", "label")
        etypes = rd.index("elements/index", "_type")
        elements = []
        for i in range(len(elabels)):
            ty = int(etypes[i])
            lab = elabels[i].decode()
            f0 = fracs[int(rd.get(EP, "frac0", i))]
            nc, ni = int(rd.get(EP, "ncoef", i)), int(rd.get(EP, "nint", i))
            if ty == INTERSECTION:
                e = Intersection(lab, rd.get(EP, "endpoints0", i), rd.get(EP, "endpoints1", i), f0,
                                 fracs[int(rd.get(EP, "frac1", i))], nc, ni)
            elif ty == BOUNDING_CIRCLE:
                e = BoundingCircle(lab, float(rd.get(EP, "radius", i)), f0, nc, ni)
            elif ty == WELL:
                e = Well(lab, float(rd.get(EP, "radius", i)), complex(rd.get(EP, "center", i)),
                         float(rd.get(EP, "head", i)), f0)
                e.phi = float(rd.get(EP, "phi", i))
            elif ty == CONST_HEAD_LINE:
                e = ConstantHeadLine(lab, rd.get(EP, "endpoints0", i), float(rd.get(EP, "head", i)), f0, nc, ni)
                e.phi = float(rd.get(EP, "phi", i))
            elif ty == IMP_CIRCLE:
                e = ImpermeableCircle(lab, float(rd.get(EP, "radius", i)), complex(rd.get(EP, "center", i)), f0,
                                      nc, ni)
            elif ty == IMP_LINE:
                key = "endpoints0" if rd.has(EP, "endpoints0", i) else "focis"
                e = ImpermeableLine(lab, rd.get(EP, key, i), f0, nc, ni)
            else:
                raise ValueError(f"element {i}: unknown type {ty}")
            e.ncoef, e.nint = nc, ni
            coef = np.asarray(rd.get(EP, "coef", i), dtype=np.complex128)
            e.coef = coef[:nc].copy()
            e.q = float(rd.get(EP, "q", i))
            e.error = float(rd.get(EP, "error", i))
            elements.append(e)
        # each fracture's elements in the saved order (the source order of its sweeps)
        if rd.has(FP, "elements", 0) and rd.has(FP, "nelements", 0):
            for i, f in enumerate(fracs):
                ids = np.asarray(rd.get(FP, "elements", i))[: int(rd.get(FP, "nelements", i))]
                f.elements = [elements[int(k)] for k in ids]
    finally:
        rd.close()
    dfn.fractures = []
    dfn.add_fracture(fracs)
    dfn.get_elements()
    dfn.consolidate_dfn()
    return dfn
