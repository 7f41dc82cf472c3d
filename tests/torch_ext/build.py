"""Builds tests/torch_ext/vfs_autograd (the libtorch autograd wrapper of the
differentiable substeps) in-tree against libvfs.so; run by
__graft_entry__.build().  Host compiler only (no device code)."""
from __future__ import annotations

import os
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
BUILD = HERE / "build"


def build(verbose: bool = False):
    from torch.utils import cpp_extension

    BUILD.mkdir(exist_ok=True)
    os.environ.setdefault("MAX_JOBS", "8")
    lib_dir = ROOT / "vkfluidsim_amd"
    return cpp_extension.load(
        name="vfs_autograd",
        sources=[str(HERE / "vfs_autograd.cpp")],
        extra_include_paths=[str(ROOT / "include")],
        extra_cflags=["-O2"],
        extra_ldflags=[f"-L{lib_dir}", "-lvfs", f"-Wl,-rpath,{lib_dir}"],
        build_directory=str(BUILD),
        with_cuda=False,
        verbose=verbose,
    )


def load():
    """Imports the built module without rebuilding (torch must be imported)."""
    import importlib.machinery
    import importlib.util

    import torch  # noqa: F401  (the module links against torch's libraries)

    so = BUILD / "vfs_autograd.so"
    if not so.exists():
        raise ImportError(f"{so} is not built; run __graft_entry__.build()")
    loader = importlib.machinery.ExtensionFileLoader("vfs_autograd", str(so))
    spec = importlib.util.spec_from_loader("vfs_autograd", loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    return mod


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
