"""The reference's options for the force path and its time step, read as PETSc reads them.

BioFEM-IA takes its parameters from PETSc's options database: `control.dat` in the working
directory as the options file, then the command line, then `<-in_dir>/control.dat`
inserted last (src/main.c:175-188), queried by name (src/main.c:190-252,
src/active_strain.c:94-106).  PETSc is not a dependency here; this module restates the
parts of its options database those calls use:

- an options string is a sequence of `-name [value]` tokens; a token is a name when it is
  `-` or `--` followed by a letter, so `-1e-3` is a value; a name followed by another name
  (or by nothing) is set without a value (a true flag);
- in files, `#` starts a comment that runs to the end of the line;
- names compare without regard to case, and setting a name again replaces its value
  (PetscOptionsSetValue): the `-in_dir` file, inserted after the command line, wins over
  it too, whatever the comment at src/main.c:177-179 says.

`act_params` fills `ActParams` from the database with the reference's names and its
defaults (the globals of src/main.c:22-64 and src/active_strain.c:47-50 are zero unless
set); like the reference, the force needs -num_gaussian_quad_points in 1..5 and
-C33_subitr_nums >= 1, which have no usable default (both 0).
"""
from __future__ import annotations

import os
import re

from .fem import ActParams

_TRUE = {"", "1", "true", "yes", "on"}
_FALSE = {"0", "false", "no", "off"}


def _is_name(tok: str) -> bool:
    return re.match(r"^--?[A-Za-z]", tok) is not None


def tokenize(text: str):
    """(name, value or None) pairs of an options string, in order."""
    toks = text.split()
    out = []
    i = 0
    while i < len(toks):
        t = toks[i]
        if not _is_name(t):
            raise ValueError(f"options: value {t!r} without a name")
        name = t.lstrip("-").lower()
        if i + 1 < len(toks) and not _is_name(toks[i + 1]):
            out.append((name, toks[i + 1]))
            i += 2
        else:
            out.append((name, None))
            i += 1
    return out


def _strip_comments(text: str) -> str:
    return "\n".join(line.split("#", 1)[0] for line in text.splitlines())


class Options:
    """An options database: names without the leading dash, lower case."""

    def __init__(self):
        self._v: dict = {}

    def insert_string(self, text: str) -> None:
        for name, value in tokenize(text):
            self._v[name] = value

    def insert_file(self, path: str, require: bool = False) -> bool:
        """PetscOptionsInsertFile: returns False for a missing optional file."""
        if not os.path.exists(path):
            if require:
                raise FileNotFoundError(path)
            return False
        with open(path) as f:
            self.insert_string(_strip_comments(f.read()))
        return True

    def insert_argv(self, argv) -> None:
        self.insert_string(" ".join(argv))

    def has(self, name: str) -> bool:
        return name.lstrip("-").lower() in self._v

    def get_string(self, name: str, default=None):
        key = name.lstrip("-").lower()
        if key not in self._v:
            return default
        v = self._v[key]
        return "" if v is None else v

    def get_int(self, name: str, default: int = 0) -> int:
        v = self.get_string(name)
        if v is None:
            return default
        for parse in (lambda t: int(t, 10), lambda t: int(t, 0)):   # decimal, then 0x...
            try:
                return parse(v)
            except ValueError:
                pass
        raise ValueError(f"option {name} needs an integer value, got {v!r}")

    def get_real(self, name: str, default: float = 0.0) -> float:
        v = self.get_string(name)
        if v is None:
            return default
        if v == "":
            raise ValueError(f"option {name} needs a real value")
        return float(v)

    def get_bool(self, name: str, default: bool = False) -> bool:
        v = self.get_string(name)
        if v is None:
            return default
        v = v.lower()
        if v in _TRUE:
            return True
        if v in _FALSE:
            return False
        raise ValueError(f"option {name}: {v!r} is not a truth value")


def load(argv=(), cwd: str = ".") -> Options:
    """The reference's database: `cwd/control.dat`, the command line, then
    `<-in_dir>/control.dat` when it exists (src/main.c:175-188)."""
    opts = Options()
    opts.insert_file(os.path.join(cwd, "control.dat"))
    opts.insert_argv(list(argv))
    in_dir = opts.get_string("-in_dir", "")
    if in_dir:
        d = in_dir if os.path.isabs(in_dir) else os.path.join(cwd, in_dir)
        opts.insert_file(os.path.join(d, "control.dat"))
    return opts


def act_params(opts: Options) -> ActParams:
    """ActParams from the database: GetUserActParams (src/active_strain.c:94-106) and the
    globals the force reads (-mu, -h0, -dt, -tisteps, -tistart, -manufactured;
    src/main.c:192-222).  -mu is the shear modulus on this path (src/active_strain.c:108).
    -gamma_pulse_t0 / -gamma_pulse_sigma set the library's ramp type 4 (a Gaussian pulse,
    no reference counterpart)."""
    n_qp = opts.get_int("-num_gaussian_quad_points", 0)
    iters = opts.get_int("-C33_subitr_nums", 0)
    if not 1 <= n_qp <= 5:
        raise ValueError(f"-num_gaussian_quad_points must be 1..5 (got {n_qp}; the reference "
                         "defaults it to 0, which it cannot run)")
    if iters < 1:
        raise ValueError(f"-C33_subitr_nums must be >= 1 (got {iters}; the reference defaults "
                         "it to 0, which leaves the stress unset)")
    return ActParams(
        n_qp=n_qp, c33_iters=iters,
        cart_fib_act=opts.get_int("-cart_fib_act", 0),
        manufactured=opts.get_int("-manufactured", 0),
        h0=opts.get_real("-h0", 0.0),
        mu=opts.get_real("-mu", 0.0),
        K=opts.get_real("-bulk_modulus", 0.0),
        gamma=opts.get_real("-muscle_act_gamma", 0.0),
        ramp_type=opts.get_int("-gamma_ramp_type", 0),
        gamma_T=opts.get_real("-gamma_T", 0.0),
        tau_up=opts.get_real("-gamma_tau_up", 0.0),
        tau_down=opts.get_real("-gamma_tau_down", 0.0),
        apd=opts.get_real("-gamma_apd", 0.0),
        pulse_t0=opts.get_real("-gamma_pulse_t0", 0.0),
        pulse_sigma=opts.get_real("-gamma_pulse_sigma", 1.0),
        manufactured_gamma0=opts.get_real("-manufactured_gamma0", 0.0),
        manufactured_T=opts.get_real("-manufactured_T", 1.0),
        ti=opts.get_int("-tistart", 0),
        tisteps=opts.get_int("-tisteps", 1),
        dt=opts.get_real("-dt", 0.0),
    )


def run_options(opts: Options) -> dict:
    """The time loop's options with the reference's defaults (src/main.c:22-64,
    192-215): density, step, Newmark and damping settings, case directory and the
    coordinate scales Input divides by."""
    return dict(
        rho=opts.get_real("-rho", 0.0), dt=opts.get_real("-dt", 0.0),
        tisteps=opts.get_int("-tisteps", 1), tistart=opts.get_int("-tistart", 0),
        timeinteg=opts.get_int("-timeinteg", 0), explicit=opts.get_int("-explicit", 0),
        damping=opts.get_int("-damping", 0), dampfactor=opts.get_real("-dampfactor", 0.0),
        newmark_gamma=opts.get_real("-newmark_gamma", 0.5),
        disable_fdyn=opts.get_int("-disable_fdyn", 0), tiout=opts.get_int("-tiout", 0),
        in_dir=opts.get_string("-in_dir", ""),
        char_length=(opts.get_real("-char_length_x", 1.0), opts.get_real("-char_length_y", 1.0),
                     opts.get_real("-char_length_z", 1.0)),
    )
