"""Renderer side of the render-buffer handoff test (test_gpu_render.py): run
as its own process after the producer has synchronized its copy, it maps the
render buffers of that process from the handle bytes in argv[1] and saves
what it reads to the .npz in argv[2]."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import vkfluidsim_amd as V  # noqa: E402


def main() -> int:
    handles = Path(sys.argv[1]).read_bytes()
    rb = V.RenderBuffers.open(handles)
    x, v, rho, a = rb.read()
    try:
        rb.export()
        refused = False
    except V.VfsError:
        refused = True
    rb.close()
    np.savez(sys.argv[2], x=x, v=v, rho=rho, a=a if a is not None else np.zeros(0, np.float32),
             export_refused=np.array(refused))
    print(f"consumer: mapped {len(x)} particles", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
