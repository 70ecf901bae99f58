"""The reference's options database as BioFEM-IA reads it (biofem_amd.options): PETSc's
token rules, comments in files, replacement by later insertions, the order of
src/main.c:175-188 (cwd control.dat, command line, then <-in_dir>/control.dat), and the
mapping onto the force path's parameters with the reference's names and defaults."""
import os

import pytest

from biofem_amd import options as O
from biofem_amd.fem import ActParams


def test_tokens_names_values_and_flags():
    toks = O.tokenize("-h0 0.01 -mu 0.3846 -twod -dt -1e-3 --tisteps 5 -x -.5")
    assert toks == [("h0", "0.01"), ("mu", "0.3846"), ("twod", None), ("dt", "-1e-3"),
                    ("tisteps", "5"), ("x", "-.5")]
    with pytest.raises(ValueError):
        O.tokenize("0.5 -h0 1")


def test_file_comments_case_and_replacement(tmp_path):
    f = tmp_path / "control.dat"
    f.write_text("# material\n-h0 0.01   -MU 0.5  # shear modulus\n\n-cart_fib_act 1\n-h0 0.02\n-bending\n")
    o = O.Options()
    assert o.insert_file(str(f))
    assert o.get_real("-h0") == 0.02             # the later value replaces the first
    assert o.get_real("-mu") == 0.5              # names ignore case
    assert o.get_int("-cart_fib_act") == 1
    assert o.get_bool("-bending") is True and o.get_bool("-membrane") is False
    assert o.get_real("-dt", 7.0) == 7.0
    assert not o.insert_file(str(tmp_path / "missing.dat"))
    with pytest.raises(FileNotFoundError):
        o.insert_file(str(tmp_path / "missing.dat"), require=True)
    with pytest.raises(ValueError):
        o.get_int("-h0")


def test_load_order_of_main(tmp_path):
    (tmp_path / "control.dat").write_text("-h0 1 -dt 2 -rho 3\n")
    case = tmp_path / "case"
    case.mkdir()
    (case / "control.dat").write_text("-dt 5 -tisteps 9\n")
    o = O.load(["-dt", "4", "-rho", "6", "-in_dir", "case"], cwd=str(tmp_path))
    assert o.get_real("-h0") == 1.0      # cwd file only
    assert o.get_real("-rho") == 6.0     # command line over the cwd file
    assert o.get_real("-dt") == 5.0      # the -in_dir file is inserted last
    assert o.get_int("-tisteps") == 9
    run = O.run_options(o)
    assert run["in_dir"] == "case" and run["newmark_gamma"] == 0.5 and run["char_length"] == (1.0, 1.0, 1.0)


def test_act_params_from_reference_names():
    o = O.Options()
    o.insert_string("-num_gaussian_quad_points 3 -C33_subitr_nums 4 -cart_fib_act 1 -h0 0.02 "
                    "-mu 0.3846 -bulk_modulus 0.8333 -muscle_act_gamma 0.12 -gamma_ramp_type 3 "
                    "-gamma_T 0.8 -gamma_tau_up 0.05 -gamma_tau_down 0.1 -gamma_apd 0.3 "
                    "-dt 1e-3 -tisteps 200 -tistart 10 -manufactured 0")
    p = O.act_params(o)
    assert p == ActParams(n_qp=3, c33_iters=4, cart_fib_act=1, manufactured=0, h0=0.02,
                          mu=0.3846, K=0.8333, gamma=0.12, ramp_type=3, gamma_T=0.8,
                          tau_up=0.05, tau_down=0.1, apd=0.3, pulse_t0=0.0, pulse_sigma=1.0,
                          manufactured_gamma0=0.0, manufactured_T=1.0, ti=10, tisteps=200,
                          dt=1e-3)


@pytest.mark.parametrize("text", ["-C33_subitr_nums 3", "-num_gaussian_quad_points 6 -C33_subitr_nums 3",
                                  "-num_gaussian_quad_points 3"])
def test_act_params_needs_quadrature_and_iterations(text):
    o = O.Options()
    o.insert_string(text)
    with pytest.raises(ValueError):
        O.act_params(o)
