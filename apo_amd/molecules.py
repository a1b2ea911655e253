"""Real-molecule inputs: MMTF structures and the ball-dataset text format.

The reference loads proteins with chemfiles (app/common/samples.cpp:10-61,
apo::loadProtein): every atom of the first frame, in residue order, at its
position, with the element's van der Waals radius or 1 when the element has
none.  chemfiles builds that frame from the file's first biological
assembly: the atom counts the reference's benchmark lists give for its
samples (app/benchmarks/benchmark_stages.cu:38-57, e.g. 7LER 158430 = 6
copies of 26405 atoms) are those of the assembly.  chemfiles is not
available here; this module decodes MMTF itself (msgpack container, the
binary codecs of the MMTF 1.0 specification) and uses Bondi's van der Waals
radii (A. Bondi, J. Phys. Chem. 68 (1964) 441), the table chemfiles ships
for the common elements.

parse_ball_dataset follows apo::parseFromDataset (samples.cpp:87-127): a
first line with the site count, then one line per site whose 2nd to 5th
whitespace-separated fields are x, y, z, r, each read as a C float (strtof,
as std::stof does).

joggle follows apo::joggle (apo/src/core/utils.cpp:14-25): every
coordinate and radius moves by pcg(4i + j) * 2a - a, the perturbation the
reference's apps apply before a build to reach general position.

Sites come back as (n, 4) float32 (x, y, z, r), the input of AlgorithmGPU.
"""
from __future__ import annotations

import ctypes
import struct

import numpy as np

from .datasets import pcg_word

# Bondi (1964) van der Waals radii in Angstrom
VDW_RADII = {
    "H": 1.20, "He": 1.40, "Li": 1.82, "C": 1.70, "N": 1.55, "O": 1.52, "F": 1.47, "Ne": 1.54,
    "Na": 2.27, "Mg": 1.73, "Si": 2.10, "P": 1.80, "S": 1.80, "Cl": 1.75, "Ar": 1.88, "K": 2.75,
    "Ni": 1.63, "Cu": 1.40, "Zn": 1.39, "Ga": 1.87, "As": 1.85, "Se": 1.90, "Br": 1.85, "Kr": 2.02,
    "Pd": 1.63, "Ag": 1.72, "Cd": 1.58, "In": 1.93, "Sn": 2.17, "Te": 2.06, "I": 1.98, "Xe": 2.16,
    "Pt": 1.72, "Au": 1.66, "Hg": 1.55, "Tl": 1.96, "Pb": 2.02, "U": 1.86,
}
# radius of an element without one (samples.cpp:55)
DEFAULT_RADIUS = 1.0


def vdw_radius(element: str) -> float:
    e = element.strip()
    e = e[:1].upper() + e[1:].lower()
    return VDW_RADII.get(e, DEFAULT_RADIUS)


# --- MMTF binary codecs -----------------------------------------------------
def _run_length(values: np.ndarray) -> np.ndarray:
    """(value, count) pairs -> the expanded sequence."""
    v = values.reshape(-1, 2)
    return np.repeat(v[:, 0], v[:, 1])


def _recursive_index(values: np.ndarray, lo: int, hi: int) -> np.ndarray:
    """Integers split into runs of the type's extreme values: each output is
    the sum of a run of lo / hi entries and the first entry after it."""
    out = []
    acc = 0
    for x in values.tolist():
        acc += x
        if x != lo and x != hi:
            out.append(acc)
            acc = 0
    return np.asarray(out, np.int64)


def decode(field):
    """One MMTF field: plain msgpack values pass through; binary fields carry
    a 12-byte big-endian header (codec, length, parameter) and the payload."""
    if not isinstance(field, (bytes, bytearray)):
        return field
    codec, length, param = struct.unpack(">iii", field[:12])
    data = field[12:]
    if codec == 1:
        return np.frombuffer(data, ">f4").astype(np.float32)
    if codec == 2:
        return np.frombuffer(data, ">i1").astype(np.int32)
    if codec == 3:
        return np.frombuffer(data, ">i2").astype(np.int32)
    if codec == 4:
        return np.frombuffer(data, ">i4").astype(np.int32)
    if codec == 5:
        return [data[i * param:(i + 1) * param].rstrip(b"\x00").decode() for i in range(length)]
    if codec == 6:
        return [chr(c) if c else "" for c in _run_length(np.frombuffer(data, ">i4")).tolist()]
    if codec == 7:
        return _run_length(np.frombuffer(data, ">i4").astype(np.int64))
    if codec == 8:
        return np.cumsum(_run_length(np.frombuffer(data, ">i4").astype(np.int64)))
    if codec == 9:
        return (_run_length(np.frombuffer(data, ">i4").astype(np.int64)) / param).astype(np.float32)
    if codec == 10:
        ints = _recursive_index(np.frombuffer(data, ">i2").astype(np.int64), -32768, 32767)
        return (np.cumsum(ints) / param).astype(np.float32)
    if codec == 11:
        return (np.frombuffer(data, ">i2").astype(np.int64) / param).astype(np.float32)
    if codec == 12:
        return (_recursive_index(np.frombuffer(data, ">i2").astype(np.int64), -32768, 32767) / param).astype(np.float32)
    if codec == 13:
        return (_recursive_index(np.frombuffer(data, ">i1").astype(np.int64), -128, 127) / param).astype(np.float32)
    if codec == 14:
        return _recursive_index(np.frombuffer(data, ">i2").astype(np.int64), -32768, 32767)
    if codec == 15:
        return _recursive_index(np.frombuffer(data, ">i1").astype(np.int64), -128, 127)
    raise ValueError(f"unknown MMTF codec {codec}")


def load_mmtf(path: str, assembly: bool = True) -> np.ndarray:
    """Atoms of the first model of an MMTF file as (n, 4) float32 sites.
    assembly: expand the first biological assembly (each transform, in
    order, applied to its chains, in order; the 4x4 matrices are row-major
    with the translation in the last column), as chemfiles does."""
    import msgpack
    with open(path, "rb") as f:
        d = msgpack.unpackb(f.read(), raw=False)
    x, y, z = (np.asarray(decode(d[k]), np.float64) for k in ("xCoordList", "yCoordList", "zCoordList"))
    group_types = np.asarray(decode(d["groupTypeList"]), np.int64)
    groups = d["groupList"]
    chains_model0 = int(d["chainsPerModel"][0])
    # atom range and radii of every chain of the first model
    chain_ranges, radii = [], []
    g = 0
    for c in range(chains_model0):
        start = len(radii)
        for _ in range(int(d["groupsPerChain"][c])):
            radii.extend(vdw_radius(e) for e in groups[int(group_types[g])]["elementList"])
            g += 1
        chain_ranges.append((start, len(radii)))
    n = len(radii)
    if n > len(x):
        raise ValueError(f"{path}: groups hold {n} atoms, coordinates {len(x)}")
    xyz = np.stack([x[:n], y[:n], z[:n]], axis=1)
    r = np.asarray(radii, np.float64)
    assemblies = d.get("bioAssemblyList") or []
    if not assembly or not assemblies:
        parts = [(xyz, r)]
    else:
        parts = []
        for t in assemblies[0]["transformList"]:
            m = np.asarray(t["matrix"], np.float64).reshape(4, 4)
            for c in t["chainIndexList"]:
                if c >= chains_model0:
                    continue
                a, b = chain_ranges[c]
                parts.append((xyz[a:b] @ m[:3, :3].T + m[:3, 3], r[a:b]))
    out = np.concatenate([np.concatenate([p, q[:, None]], axis=1) for p, q in parts]).astype(np.float32)
    return out


_strtof = None


def _to_float32(token: str) -> float:
    """A decimal string rounded once to FP32, as strtof / std::stof do."""
    global _strtof
    if _strtof is None:
        _strtof = ctypes.CDLL(None).strtof
        _strtof.restype = ctypes.c_float
        _strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    return _strtof(token.encode(), None)


def parse_ball_dataset(path: str) -> np.ndarray:
    """Sites of a ball-dataset text file (apo::parseFromDataset)."""
    with open(path) as f:
        count = int(f.readline().split()[0])
        rows = []
        for line in f:
            tok = line.split()
            if len(tok) < 5:
                continue
            rows.append([_to_float32(t) for t in tok[1:5]])
    out = np.asarray(rows, np.float32).reshape(-1, 4)
    if len(out) != count:
        raise ValueError(f"{path}: header says {count} sites, found {len(out)}")
    return out


def joggle(sites: np.ndarray, amount: float) -> np.ndarray:
    """apo::joggle on FP32 sites: each value plus pcg(4i + j) * 2 * amount -
    amount, in FP64 (pcg = word * 2^-32, utils.cpp:6-12), rounded to FP32."""
    s = np.asarray(sites, np.float32).reshape(-1, 4)
    v = np.arange(s.size, dtype=np.uint64).astype(np.uint32)
    noise = np.ldexp(pcg_word(v).astype(np.float64), -32) * (2.0 * amount) - amount
    return (s.astype(np.float64).ravel() + noise).astype(np.float32).reshape(-1, 4)
