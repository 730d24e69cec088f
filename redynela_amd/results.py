"""Result files in the reference's binary layout.

The reference saves a structure at each save time with io_Data::structure_write
(FE/io_Data.C:487-510) into `<deck base>.NNN.bin`, NNN counting from 001
(FE/io_Structure.C:60-68, incremental files switched on by
Applications/DynELA_solve/main.C:104).  The bytes are raw native doubles:

  Structure::writeData (FE/Structure.C:936-953)
    version Vec3D (VERSION, RELEASE, UNDER_RELEASE) = (3, 1, 0), then the
    check Vec3D (1.02, ALPHA_VERSION, 0)         FE/Structure.C:989-1011,
                                                 FE/Structure.h:34-35, Const.h:43-46
    current time
    per domain, Domain::writeData (FE/Domain.C:1407-1431):
      version + check, current time
      per node, Node::write (FE/Node.C:156-162): coords(3), then the Current
      NodalField (FE/NodalField.C:102-117): ro, dro, e, de, T, mat_speed(3),
      dmat_speed(3), fe(3), delta_disp(3), disp(3)
      separator (69, ALPHA_VERSION, 0)
      per element, per integration point, IntegrationPoint::write
      (FE/IntegrationPoint.C:155-168): p, evp, S(9), Sig(9), Eps(9), dEps(9),
      EpsPlas(9), each tensor row-major (LA/Tensor2.C:784-788)
      separator
    separator after each domain

The Current NodalField of a node is one of two that the solver swaps at the
end of every increment (FE/Node.C:219-228 from FE/ExplicitSolverCH.C:467).
Domain::createElement sets ro = density and T = the material's initial
temperature in the field that is Current before the first increment
(FE/Domain.C:1100-1104); the other field keeps ro = T = 0 (NodalField::Clear,
FE/NodalField.C:43-66), and the explicit Chung-Hulbert path writes neither.
So the written ro and T are (density, T0) after an even number of
increments and (0, 0) after an odd one; dro, e, de and the nodal external
force fe stay 0 on this path (no force boundary conditions).
"""
import os

import numpy as np

from ._lib import DnlgError

VERSION = (3.0, 1.0, 0.0)
VERSION_CHECK = 1.02
SEPARATOR = 69.0
ALPHA_VERSION = 0.0

NODE_DOUBLES = 23       # coords 3, ro, dro, e, de, T, v 3, a 3, fe 3, du 3, d 3
IP_DOUBLES = 47         # p, evp, S 9, Sig 9, Eps 9, dEps 9, EpsPlas 9
_IDENT = np.eye(3).reshape(9)


def _header(t):
    return np.array(VERSION + (VERSION_CHECK, ALPHA_VERSION, 0.0, t))


def _sep():
    return np.array((SEPARATOR, ALPHA_VERSION, 0.0))


def node_block(coords, v, a, du, d, increment, density, T0=0.0, fe=None):
    """[nn][23] Node::write records; the Current field's ro and T follow the
    swap parity of `increment` (the number of completed increments): the
    `density` and `T0` given (scalars or per node: after a restart, the
    restored file's values) after an even number, 0 after an odd one."""
    nn = coords.shape[0]
    rec = np.zeros((nn, NODE_DOUBLES))
    rec[:, 0:3] = coords
    if increment % 2 == 0:
        rec[:, 3] = density
        rec[:, 7] = T0
    rec[:, 8:11] = v
    rec[:, 11:14] = a
    if fe is not None:
        rec[:, 14:17] = fe
    rec[:, 17:20] = du
    rec[:, 20:23] = d
    return rec


def ip_block(p, evp, S, eps, deps, epsplas):
    """[n_ip][47] IntegrationPoint::write records.  Tensors are [n_ip][9]
    row-major; Sig = S + p I as FE/Element.C:400-401 forms it."""
    n = evp.shape[0]
    rec = np.empty((n, IP_DOUBLES))
    rec[:, 0] = p
    rec[:, 1] = evp
    rec[:, 2:11] = S
    rec[:, 11:20] = S + p[:, None] * _IDENT[None, :]
    rec[:, 20:29] = eps
    rec[:, 29:38] = deps
    rec[:, 38:47] = epsplas
    return rec


def write(path, t, nodes, ips):
    """Write one structure with one domain; nodes/ips from node_block/ip_block."""
    parts = [_header(t), _header(t), nodes.reshape(-1), _sep(), ips.reshape(-1), _sep(), _sep()]
    with open(path, "wb") as f:
        for p in parts:
            f.write(np.ascontiguousarray(p, dtype="<f8").tobytes())


def read(path, n_nodes, n_ips):
    """Read a one-domain result file; checks the version marks and every
    separator as Structure::checkBinaryVersion does (FE/Structure.C:956-987)
    and returns dict(time, domain_time, nodes [nn][23], ips [n_ip][47])."""
    raw = np.fromfile(path, dtype="<f8")
    want = 7 + 7 + n_nodes * NODE_DOUBLES + 3 + n_ips * IP_DOUBLES + 3 + 3
    if raw.size != want:
        raise DnlgError("%s: %d doubles, a structure of %d nodes and %d points has %d"
                        % (path, raw.size, n_nodes, n_ips, want))
    if raw[3] != VERSION_CHECK or raw[10] != VERSION_CHECK:
        raise DnlgError("%s: this recording format is not compatible (check %g)" % (path, raw[3]))
    o = 14
    nodes = raw[o:o + n_nodes * NODE_DOUBLES].reshape(n_nodes, NODE_DOUBLES)
    o += n_nodes * NODE_DOUBLES
    seps = [o]
    o += 3
    ips = raw[o:o + n_ips * IP_DOUBLES].reshape(n_ips, IP_DOUBLES)
    o += n_ips * IP_DOUBLES
    seps += [o, o + 3]
    for s in seps:
        if raw[s] != SEPARATOR:
            raise DnlgError("%s: may be file is truncated (separator at double %d)" % (path, s))
    return dict(time=raw[6], domain_time=raw[13], nodes=nodes, ips=ips)


def restore(dom, path):
    """Structure::readData of a restart (FE/Structure.C:915-931,
    FE/Domain.C:1378-1404, as Sources/Applications/DynELA_solve/main.C:106-130
    drives it) into a GPU Domain (FP64 form) that has not begun: coordinates,
    the Current nodal fields (v, a, Delta u, d), the point state (p, evp, S,
    and Eps, EpsPlas when strains are tracked) and the current time.  The next
    solve recomputes the mass from the restored coordinates and starts a
    fresh time-step history with increment 0, as the restarted reference run
    does.  Returns the file's time."""
    from . import _lib as L
    nn = dom.n_nodes
    nip = 8 if dom.dim == 3 else 4
    nipt = dom.field(L.F_EVP).size
    rec = read(path, nn, nipt)
    nodes, ips = rec["nodes"], rec["ips"]
    for f, c in ((L.F_COORDS, 0), (L.F_V, 8), (L.F_A, 11), (L.F_DU, 17), (L.F_D, 20)):
        dom.set_field(f, nodes[:, c:c + 3])
    dom.set_field(L.F_EVP, ips[:, 1])
    dom.set_field(L.F_S, ips[:, 2:11].T)
    dom.set_field(L.F_P, ips[:, 0].reshape(-1, nip)[:, 0])  # equal at the points of an element
    if L.lib().dnlg_field_size(dom.h, L.F_EPS) > 0:
        dom.set_field(L.F_EPS, ips[:, 20:29].T)
        dom.set_field(L.F_EPSPLAS, ips[:, 38:47].T)
    dom.set_time(rec["time"])
    # Node::read puts the file's ro and T into the Current field (FE/Node.C:148);
    # the other field stays cleared, so from here on the files carry these
    # values after an even number of increments and 0 after an odd number
    dom.restored_density_T = (nodes[:, 3].copy(), nodes[:, 7].copy())
    return rec["time"]


def _tensors(dom, f, nipt):
    return dom.field(f).reshape(9, nipt).T


def domain_blocks(dom, density, T0=0.0):
    """node and integration-point records of a GPU Domain (strain tracking on)."""
    from . import _lib as L
    tm = dom.time()
    nn = dom.coords.shape[0]
    evp = dom.evp
    nipt = evp.size
    if L.lib().dnlg_field_size(dom.h, L.F_DEPS) <= 0:
        raise DnlgError("result files need strain tracking (Domain(..., track_strain=True))")
    vec = lambda f: dom.field(f).reshape(nn, 3)
    nodes = node_block(dom.coords, vec(L.F_V), vec(L.F_A), vec(L.F_DU), vec(L.F_D),
                       tm["increment"], density, T0)
    p = np.repeat(dom.field(L.F_P), nipt // dom.field(L.F_P).size)
    ips = ip_block(p, evp, _tensors(dom, L.F_S, nipt), _tensors(dom, L.F_EPS, nipt),
                   _tensors(dom, L.F_DEPS, nipt), _tensors(dom, L.F_EPSPLAS, nipt))
    return tm["t"], nodes, ips


class ResultFiles:
    """Incremental result files `<base>.NNN.bin` (FE/io_Structure.C:60-68)."""

    def __init__(self, base, density, T0=0.0):
        self.base, self.density, self.T0 = base, density, T0
        self.count = 1
        self.paths = []

    def __call__(self, t, dom):
        tt, nodes, ips = domain_blocks(dom, self.density, self.T0)
        path = "%s.%03d.bin" % (self.base, self.count)
        self.count += 1
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        write(path, t, nodes, ips)
        self.paths.append(path)
        return path


PLOT_HEADER = "#DynELA_plot : Time history of computation\n#plotted : Used_timeStep timeStep kinetic_energy\n"


class PlotFile:
    """The domain's `.plot` history file: opened with its header by the first
    Domain::solve (FE/Domain.C:1244-1255), then one line per increment,
    "%8.4E %8.4E %8.4E %8.4E" of time, time step, computed time step and
    kinetic energy (FE/ExplicitSolverCH.C:459-461).  Its name is the deck's
    base name, '_', the domain name (defaultDomain, FE/Domain.C:55) and
    '.plot'."""

    def __init__(self, path):
        self.path = path
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(path, "w") as f:
            f.write(PLOT_HEADER)

    def append(self, rows):
        with open(self.path, "a") as f:
            for r in rows:
                f.write("%8.4E %8.4E %8.4E %8.4E\n" % (r[0], r[1], r[2], r[3]))


def read_plot(path):
    """[n][4] values of a .plot file (header checked)."""
    with open(path) as f:
        text = f.read()
    if not text.startswith(PLOT_HEADER):
        raise DnlgError("%s: not a DynELA .plot file" % path)
    rows = [list(map(float, ln.split())) for ln in text[len(PLOT_HEADER):].splitlines() if ln.strip()]
    return np.array(rows).reshape(-1, 4)


class HistoryFileWriter:
    """A HistoryFile's text file (FE/HistoryFile.C:114-141): header
    "#DynELA_plot history file" and "#plotted : " with the column names, then
    per saved increment "%10.7E " of the time and of each value."""

    def __init__(self, path, columns):
        self.path = path
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(path, "w") as f:
            f.write("#DynELA_plot history file\n#plotted : " + "".join(c + " " for c in columns) + "\n")

    def append(self, rows):
        with open(self.path, "a") as f:
            for r in rows:
                f.write("".join("%10.7E " % v for v in r) + "\n")
