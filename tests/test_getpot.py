"""CPU: the GetPot reader of the C++ layer (include/tk/getpot.hxx) on the reference's own
parameter file (examples/data, kept as tests/golden/rsfe_data.pot) and on the input rules
GetPot documents (sections, relative sections, comments, 'name = value' spellings, quoted
values, numbers in float notation for ints, redefinition, missing names and files)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "getpot_dump.cpp")


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("gp") / "getpot_dump")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", SRC, "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

    def run(path, prefix, *names):
        out = subprocess.run([exe, path, prefix, *names], capture_output=True, text=True,
                             check=True).stdout
        var, get = {}, {}
        for line in out.splitlines():
            f = line.split("\t")
            if f[0] == "var":
                var[f[1]] = f[2] if len(f) > 2 else ""
            else:
                get[f[1]] = (int(f[2]), float(f[3]), f[4] if len(f) > 4 else "")
        return var, get
    return run


def test_reference_parameter_file(dump):
    """examples/data: the values its init section sets (rsfe_splitbregman_imp.hxx:582-796
    reads them under splitbregman/init/); the commented ones (nu, alpha, everything in
    onthego) are absent"""
    names = ["verbosity", "maxiter", "showfrequency", "dumpfrequency", "outputname", "sigma",
             "lamda", "lamda1", "lamda2", "beta", "dt", "ls_steps", "epsilon", "a0", "b0",
             "auto_extract_conn_comp", "cc_init_pixel_x", "cc_init_pixel_y", "cc_init_pixel_z",
             "cc_binary_output", "cc_init_variables", "auto_extract_conn_comp_freq", "bc",
             "ls_tol", "edge_detector", "edge_detector_sigma", "gaussian_pixel_approach", "tol",
             "nu", "alpha"]
    var, get = dump(os.path.join(HERE, "golden", "rsfe_data.pot"), "splitbregman/init/", *names)
    ints = {"verbosity": 1, "maxiter": 10, "showfrequency": 1, "dumpfrequency": 0,
            "ls_steps": 5, "auto_extract_conn_comp": 0, "cc_binary_output": 1,
            "cc_init_variables": 1, "auto_extract_conn_comp_freq": 1, "bc": 2,
            "edge_detector": 1, "gaussian_pixel_approach": 0}
    reals = {"sigma": 3.5, "lamda": 1e-2, "lamda1": 1e-5, "lamda2": 1e-5, "beta": 100.0,
             "dt": 1.0, "epsilon": 1e-3, "a0": 0.0, "b0": 1.0, "cc_init_pixel_x": 60.0,
             "cc_init_pixel_y": 28.0, "cc_init_pixel_z": 21.0, "ls_tol": 0.0,
             "edge_detector_sigma": 0.0, "tol": 1e-5}
    for k, v in ints.items():
        assert get[k][0] == v, k
    for k, v in reals.items():
        assert get[k][1] == v, k
    assert get["outputname"][2] == "result_"
    assert get["nu"] == (-1, -1.0, "") and get["alpha"] == (-1, -1.0, "")
    assert not any(k.startswith("splitbregman/onthego/") for k in var)
    assert all(k.startswith("splitbregman/init/") for k in var)


def test_getpot_rules(dump, tmp_path):
    p = str(tmp_path / "rules.pot")
    with open(p, "w") as f:
        f.write("top = 1\n"
                "[a]\n x=2   # trailing comment\n"
                "  [./b]\n  y =3\n  z= 4\n  w = 2.0e1\n  s = 'two words'\n  y = 5\n"
                "  [../c]\n  q = 6\n"
                "[../]\n"
                "[]\n r = 7\n"
                "[d/e]\n t = x1\n")
    var, get = dump(p, "", "top", "r")
    assert var == {"top": "1", "a/x": "2", "a/b/y": "5", "a/b/z": "4", "a/b/w": "2.0e1",
                   "a/b/s": "two words", "a/c/q": "6", "r": "7", "d/e/t": "x1"}
    _, get = dump(p, "a/b/", "y", "w", "s", "missing")
    assert get["y"][0] == 5 and get["w"][0] == 20 and get["s"][2] == "two words"
    assert get["missing"] == (-1, -1.0, "")
    _, get = dump(p, "d/e/", "t")
    assert get["t"] == (-1, -1.0, "x1")  # not a number: the default
    var, get = dump(str(tmp_path / "absent.pot"), "", "top")
    assert var == {} and get["top"] == (-1, -1.0, "")
