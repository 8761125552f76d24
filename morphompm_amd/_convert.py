"""Array-kind plumbing: numpy / torch in, same kind out; device tensors inside."""
import numpy as np
import torch


class Like:
    """Remembers the caller's array kind (numpy or torch + device) and dtype."""

    def __init__(self, ref, dtype=None):
        self.is_torch = isinstance(ref, torch.Tensor)
        self.device = ref.device if self.is_torch else None
        if dtype is not None:
            self.dtype = dtype
        elif self.is_torch:
            self.dtype = ref.dtype if ref.dtype in (torch.float32, torch.float64) else torch.float64
        else:
            self.dtype = torch.float32 if np.asarray(ref).dtype == np.float32 else torch.float64

    def out(self, t):
        if self.is_torch:
            return t.to(self.device, copy=True)
        return t.detach().cpu().numpy().copy()


def to_dev(a, dtype, device="cuda"):
    if a is None:
        return None
    if isinstance(a, torch.Tensor):
        return a.detach().to(device=device, dtype=dtype)
    return torch.as_tensor(np.asarray(a), dtype=dtype, device=device)
