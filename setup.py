"""Build backend hook: `pip install .` compiles the HIP library and the pybind11 module.

setuptools drives packaging only; the native parts are built by pygranite_amd/build.py
(hipcc --offload-arch=gfx950 for the kernels, g++ for the host code and the extension), the
same recipe as `python -m pygranite_amd.build`. The wheel holds the top-level `pygranite`
package (the reference's import name) and `pygranite_amd` with libgranite_hip.so, the
extension and the C header.
"""
import importlib.util
import os
import shutil

from setuptools import setup
from setuptools.command.build_py import build_py
from setuptools.dist import Distribution

ROOT = os.path.dirname(os.path.abspath(__file__))


class BuildNative(build_py):
    def run(self):
        spec = importlib.util.spec_from_file_location(
            "_granite_build", os.path.join(ROOT, "pygranite_amd", "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build(verbose=True)
        # the C header travels inside the package so that installed users can bind the ABI
        inc = os.path.join(ROOT, "pygranite_amd", "include")
        os.makedirs(inc, exist_ok=True)
        shutil.copy2(os.path.join(ROOT, "include", "granite_hip.h"), inc)
        super().run()


class NativeDistribution(Distribution):
    def has_ext_modules(self):  # platform wheel: it carries gfx950 code objects
        return True


setup(
    name="pygranite-mi355x",
    version="1.5.0",
    description="pygranite-compatible particle trajectory engine with hand-written gfx950 "
                "(MI355X) HIP kernels",
    license="MIT",
    python_requires=">=3.9",
    install_requires=["numpy"],
    packages=["pygranite", "pygranite_amd"],
    package_data={"pygranite_amd": ["*.so", "csrc/*", "include/*.h"]},
    cmdclass={"build_py": BuildNative},
    distclass=NativeDistribution,
    zip_safe=False,
)
