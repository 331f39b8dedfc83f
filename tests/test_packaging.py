"""The package bootstraps from a clean checkout and installs as the reference's `pygranite`.

A clean copy of the tree (the files git tracks or would track; no built libraries) must import
its pure-Python modules, refuse API use with the build hint, and `pip install` into a target
directory (building the HIP library with hipcc for gfx950), after which `import pygranite`
gives the reference's module surface (source/pygranite.cpp:3-152). CPU only: the built
library is loaded, no kernel runs.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _clean_copy(dst):
    if shutil.which("git") is None or not os.path.isdir(os.path.join(ROOT, ".git")):
        pytest.skip("needs a git checkout to list the tree's files")
    out = subprocess.run(["git", "ls-files", "-co", "--exclude-standard", "-z"], cwd=ROOT,
                         capture_output=True, check=True).stdout.decode()
    n = 0
    for rel in filter(None, out.split("\0")):
        if rel.endswith(".so"):  # ignored working folders are already left out by git
            continue
        src = os.path.join(ROOT, rel)
        if not os.path.isfile(src):
            continue
        os.makedirs(os.path.join(dst, os.path.dirname(rel)), exist_ok=True)
        shutil.copy2(src, os.path.join(dst, rel))
        n += 1
    assert n > 20
    return dst


def _py(code, cwd, extra_path=None):
    env = dict(os.environ)
    env["PYTHONPATH"] = extra_path or ""
    return subprocess.run([sys.executable, "-c", code], cwd=cwd, env=env, capture_output=True,
                          text=True, timeout=300)


@pytest.fixture(scope="module")
def clean_tree(tmp_path_factory):
    return _clean_copy(str(tmp_path_factory.mktemp("clean") / "tree"))


def test_unbuilt_tree_imports_build_and_refuses_api(clean_tree):
    assert not os.path.exists(os.path.join(clean_tree, "pygranite_amd", "libgranite_hip.so"))
    r = _py("import pygranite_amd, pygranite_amd.build, pygranite_amd.synthetic, pygranite_amd.capi\n"
            "from pygranite_amd import build, serving\n"
            "assert callable(build.build)\n"
            "import __graft_entry__\n"
            "try:\n"
            "    pygranite_amd.IntegratorSettings\n"
            "except ImportError as e:\n"
            "    assert 'pygranite_amd.build' in str(e), e\n"
            "    print('refused')\n", clean_tree, clean_tree)
    assert r.returncode == 0, r.stderr
    assert "refused" in r.stdout


def test_pip_install_gives_top_level_pygranite(clean_tree, tmp_path):
    target = str(tmp_path / "site")
    r = subprocess.run([sys.executable, "-m", "pip", "install", "--no-build-isolation", "--no-deps",
                        "--no-index", "--target", target, clean_tree],
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    for rel in ("pygranite/__init__.py", "pygranite_amd/libgranite_hip.so",
                "pygranite_amd/include/granite_hip.h"):
        assert os.path.exists(os.path.join(target, rel)), rel
    code = (
        "import pygranite as pg\n"
        "assert pg.ABI_VERSION >= 1\n"
        "assert int(pg.Space.Space3D) == 0 and int(pg.Space.Space2D) == 1\n"
        "assert int(pg.Integrator.ClassicRungeKutta) == 1\n"
        "assert int(pg.AbortReason.Wind) == 4\n"
        "s = pg.IntegratorSettings()\n"
        "assert s.DeltaT == 1.0 and s.DataTimeDistance == 60.0 and s.SaveInterval == 1\n"
        "assert list(s.GridScale) == [25.0, 25.0, 25.0]\n"
        "for name in ('TrajectoryIntegrator', 'DataLoader', 'TrajectorySet'):\n"
        "    assert hasattr(pg, name), name\n"
        "print('ok', pg.__file__)\n")
    # run from an unrelated directory so that only the installed copy is importable
    r = _py(code, str(tmp_path), target)
    assert r.returncode == 0, r.stderr
    assert "ok" in r.stdout and target in r.stdout
