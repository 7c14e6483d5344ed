"""Hypercube files for the command-line program tests: genericIO-style JSON
descriptors (n1, o1, d1, ..., filename; raw little-endian float32) and SEP
text headers (key=value, in=, data_format native_float or xdr_float)."""
import json
import os

import numpy as np


def write_json(path, vals, axes):
    """axes: [(n, o, d), ...], axis 1 fastest; vals reshaped to them."""
    desc = {}
    for k, (n, o, d) in enumerate(axes, 1):
        desc["n%d" % k], desc["o%d" % k], desc["d%d" % k] = n, o, d
    desc["filename"] = os.path.basename(path) + ".dat"
    with open(path, "w") as f:
        json.dump(desc, f)
    np.ascontiguousarray(vals, np.float32).tofile(path + ".dat")


def write_sep(path, vals, axes, big_endian=True):
    lines = ["n%d=%d o%d=%r d%d=%r" % (k, n, k, o, k, d) for k, (n, o, d) in enumerate(axes, 1)]
    lines.append('in="%s@"' % path)
    lines.append('data_format="%s"' % ("xdr_float" if big_endian else "native_float"))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    np.ascontiguousarray(vals, ">f4" if big_endian else "<f4").tofile(path + "@")


def read(path):
    """(values as float32, [(n, o, d), ...]) of a JSON or SEP descriptor."""
    text = open(path).read()
    if path.endswith(".json"):
        kv = {k: str(v) for k, v in json.loads(text).items()}
        data, fmt = kv["filename"], "native_float"
    else:
        kv = {}
        for tok in text.split():
            if "=" in tok:
                k, v = tok.split("=", 1)
                kv[k] = v.strip('"')
        data, fmt = kv["in"], kv.get("data_format", "xdr_float")
    if not os.path.isabs(data) and not os.path.exists(data):
        data = os.path.join(os.path.dirname(path), data)
    axes = []
    k = 1
    while "n%d" % k in kv:
        axes.append((int(kv["n%d" % k]), float(kv.get("o%d" % k, 0)),
                     float(kv.get("d%d" % k, 1))))
        k += 1
    vals = np.fromfile(data, ">f4" if fmt == "xdr_float" else "<f4").astype(np.float32)
    return vals, axes
