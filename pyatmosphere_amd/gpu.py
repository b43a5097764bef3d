"""Backend switch (gpu.py:4-20 of pyAtmosphere).

This package is the device implementation of the path: `config['use_gpu']`
is True and device arrays are torch tensors on the current HIP device.  The
host (numpy) path is pyAtmosphere itself; setting the switch to False makes
every device operation raise instead of silently computing on the host.
"""
import numpy as np

config = {
    "use_gpu": True,
}


def _require_gpu():
    if not config["use_gpu"]:
        raise RuntimeError("pyatmosphere_amd runs on the GPU only; the numpy path is pyAtmosphere itself "
                           "(set gpu.config['use_gpu'] = True)")


def get_xp():
    """The array module of device arrays (torch)."""
    _require_gpu()
    import torch
    return torch


def device():
    _require_gpu()
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is visible")
    return torch.device("cuda", torch.cuda.current_device())


def get_array(xp_array):
    """Device array -> numpy array (gpu.py:17-20)."""
    if hasattr(xp_array, "detach"):
        return xp_array.detach().cpu().numpy()
    return np.asarray(xp_array)
