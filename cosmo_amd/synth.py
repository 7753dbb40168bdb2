"""Seeded synthetic shell-model inputs: valence spaces, Oxbash ``.int`` files and
CoSMo system XML.

The interaction is isospin-conserving and written in the Oxbash text format read by
``ReadOxbashINT`` (include/sm/sphericalint.cxx:145-246): a first line
``-count e_1 .. e_L ACORE ASCALE SPOW`` (the minus sign asks the reader to scale every
TBME by (ASCALE/(Av+ACORE))**SPOW, sphericalint.cxx:208-215), then one line per TBME
``a b c d J T V`` with 1-based level indices.  Values are written with 4 decimals, so
the float parse of the reader (std::stof) reproduces them exactly on every platform.

The list of allowed TBMEs follows the pair enumeration of GenerateEmptyINT
(include/sm/sphericalint.cxx:420-472): pairs (a<=b), (c<=d), (ab)<=(cd), every J
allowed by both triangles, T in {0,1}, identical-orbit pairs only with J+T odd, and
parity conserved.  For the sd shell this gives 63 TBMEs and for the fp shell 195.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np

# (name, n, l, 2j) in the index order used by the workloads (XML index = position + 1)
SHELLS = {
    "sd": [("0d5", 0, 2, 5), ("1s1", 1, 0, 1), ("0d3", 0, 2, 3)],
    "fp": [("0f7", 0, 3, 7), ("1p3", 1, 1, 3), ("0f5", 0, 3, 5), ("1p1", 1, 1, 1)],
    # two oscillator shells (0p: 1 quantum, sd: 2) for truncations by quanta
    "psd": [("0p3", 0, 1, 3), ("0p1", 0, 1, 1), ("0d5", 0, 2, 5), ("1s1", 1, 0, 1), ("0d3", 0, 2, 3)],
}
CORES = {"sd": ("16O", 8, 8, 16.0, 18.0, 0.3), "fp": ("40Ca", 20, 20, 40.0, 42.0, 0.3),
         "psd": ("4He", 2, 2, 4.0, 6.0, 0.3)}
SPE_RANGE = {"sd": (-4.0, 2.0), "fp": (-2.0, 6.0), "psd": (-6.0, 4.0)}
TBME_SIGMA = {"sd": 1.5, "fp": 1.2, "psd": 1.5}


def allowed_tbmes(shell: str):
    """(a, b, c, d, J, T) with 1-based levels, in file order."""
    orbs = SHELLS[shell]
    L = len(orbs)
    pairs = [(a, b) for a in range(L) for b in range(a, L)]
    out = []
    for i, (a, b) in enumerate(pairs):
        for (c, d) in pairs[i:]:
            la, lb, lc, ld = (orbs[x][2] for x in (a, b, c, d))
            if (la + lb + lc + ld) & 1:
                continue
            ja, jb, jc, jd = (orbs[x][3] for x in (a, b, c, d))
            jmin = max(abs(ja - jb), abs(jc - jd)) // 2
            jmax = min(ja + jb, jc + jd) // 2
            for J in range(jmin, jmax + 1):
                for T in (0, 1):
                    if a == b and (J + T) % 2 == 0:
                        continue
                    if c == d and (J + T) % 2 == 0:
                        continue
                    out.append((a + 1, b + 1, c + 1, d + 1, J, T))
    return out


@dataclass
class Interaction:
    shell: str
    spe: list
    tbmes: list  # (a, b, c, d, J, T, V)
    acore: float
    ascale: float
    spow: float

    def text(self) -> str:
        lines = ["! synthetic isospin interaction (seeded), shell %s" % self.shell,
                 "! levels: " + " ".join("%d=%s" % (i + 1, o[0]) for i, o in enumerate(SHELLS[self.shell]))]
        head = "%d " % (-len(self.tbmes)) + " ".join("%.4f" % e for e in self.spe)
        head += " %.4f %.4f %.4f" % (self.acore, self.ascale, self.spow)
        lines.append(head)
        for (a, b, c, d, J, T, V) in self.tbmes:
            lines.append("%3d%3d%3d%3d %4d%3d %10.4f" % (a, b, c, d, J, T, V))
        return "\n".join(lines) + "\n"


def make_interaction(shell: str, seed: int) -> Interaction:
    rng = np.random.default_rng(seed)
    lo, hi = SPE_RANGE[shell]
    spe = [round(float(x), 4) for x in rng.uniform(lo, hi, len(SHELLS[shell]))]
    keys = allowed_tbmes(shell)
    vals = rng.normal(0.0, TBME_SIGMA[shell], len(keys))
    tb = [k + (round(float(v), 4),) for k, v in zip(keys, vals)]
    _, _, _, acore, ascale, spow = CORES[shell]
    return Interaction(shell, spe, tb, acore, ascale, spow)


def _jstr(twoj: int) -> str:
    return "%d/2" % twoj if twoj % 2 else "%d" % (twoj // 2)


def system_xml(shell: str, name: str, nv: int, zv: int, twojz: int, parity: str,
               intfile: str, mbs: str | None = None, matrix: str | None = None,
               matrix_attrs: dict | None = None, system_attrs: dict | None = None,
               system_children: str = "") -> str:
    """CoSMo system XML (layout of include/xsm/schema.h:6-62).  matrix_attrs adds
    attributes such as JJ / TT to the system's <matrix> node (written when either
    matrix or matrix_attrs is given); system_attrs adds attributes such as the J / T
    filters of the state labelling to the <system> node, system_children elements such as
    the <hwrejection> / <rejection> basis truncations."""
    core, cn, cz, acore, _, _ = CORES[shell]
    av = nv + zv
    def spin_str(twice):  # a spin as the reference's XML writes it: "1", "3/2", "-1/2"
        x = Fraction(twice, 2)
        return str(x.numerator) if x.denominator == 1 else "%d/%d" % (x.numerator, x.denominator)
    jzs = spin_str(twojz)
    tzs = spin_str(nv - zv)  # valence Tz = (N - Z) / 2
    orb = "\n".join(
        '\t\t<orbital name="%s" n="%d" l="%d" j="%s" type="pn" index="%d"/>' % (o[0], o[1], o[2], _jstr(o[3]), i + 1)
        for i, o in enumerate(SHELLS[shell]))
    sattrs = "".join(' %s="%s"' % kv for kv in (system_attrs or {}).items())
    extra = ""
    if mbs:
        extra += '\t\t<?ManyBodyStates name="%s" ?>\n' % mbs
    if matrix or matrix_attrs:
        attrs = "".join(' %s="%s"' % kv for kv in (matrix_attrs or {}).items())
        extra += '\t\t<matrix name="%s" format="pp"%s/>\n' % (matrix or name + ".HH", attrs)
    if system_children:
        extra += "".join("\t\t%s\n" % line for line in system_children.strip().splitlines())
    return (
        '<?xml version="1.0" encoding="ISO-8859-1"?>\n<cosmo>\n'
        '\t<valencespace name="%s">\n\t\t<core name="%s" N="%d" Z="%d"/>\n%s\n\t</valencespace>\n'
        '\t<hamiltonian name="synthetic">\n\t\t<file format="ascii">\n\t\t\t<name>%s</name>\n'
        '\t\t</file>\n\t</hamiltonian>\n'
        '\t<system name="%s" A="%d" Av="%d" Nv="%d" Zv="%d" Tz="%s" Jz="%s" P="%s"%s>\n%s\t</system>\n'
        '</cosmo>\n' % (shell, core, cn, cz, orb, intfile, name, cn + cz + av, av, nv, zv, tzs, jzs,
                        parity, sattrs, extra))


def write_system(dirpath: str, shell: str, name: str, nv: int, zv: int, twojz: int = 0,
                 parity: str = "+", seed: int = 0, matrix_attrs: dict | None = None,
                 system_attrs: dict | None = None, system_children: str = "") -> str:
    """Write ``<name>.int`` and ``<name>.xml`` into dirpath; return the XML path."""
    os.makedirs(dirpath, exist_ok=True)
    inter = make_interaction(shell, seed)
    with open(os.path.join(dirpath, name + ".int"), "w") as f:
        f.write(inter.text())
    xml = os.path.join(dirpath, name + ".xml")
    with open(xml, "w") as f:
        f.write(system_xml(shell, name, nv, zv, twojz, parity, name + ".int", matrix_attrs=matrix_attrs,
                           system_attrs=system_attrs, system_children=system_children))
    return xml


def random_requests(ca, count: int, seed: int, p_sd: float = 1.0, max_sd: int = 6, max_fp: int = 4,
                    dirpath: str | None = None):
    """Seeded random solve requests (the batched and the serving workloads): shell (sd with
    probability p_sd, else fp), Nv and Zv uniform in 1..cap, 2Jz uniform over the values
    with the parity of A in [-2A, 2A] (empty M-spaces are drawn again), the natural
    parity of the shell, and an interaction of its own seed.  ca is the cosmo_amd module
    (its System tells the dimension).  Returns ([(shell, nv, zv, twojz, xml, system)], rng)
    -- the generator continues for the request's other draws."""
    import tempfile
    rng = np.random.default_rng(seed)
    d = dirpath or tempfile.mkdtemp(prefix="cosmo_req_")
    out = []
    while len(out) < count:
        shell = "sd" if rng.random() < p_sd else "fp"
        cap = max_sd if shell == "sd" else max_fp
        nv, zv = int(rng.integers(1, cap + 1)), int(rng.integers(1, cap + 1))
        A = nv + zv
        jmax = 2 * A
        twojz = int(rng.choice([j for j in range(-jmax, jmax + 1) if (j - A) % 2 == 0]))
        par = "+" if shell == "sd" else ("-" if A % 2 else "+")
        xml = write_system(d, shell, "r%d" % len(out), nv, zv, twojz, par, int(rng.integers(1 << 30)))
        s = ca.System.from_xml(xml)
        if s.dim < 1:
            continue
        out.append((shell, nv, zv, twojz, xml, s))
    return out, rng
