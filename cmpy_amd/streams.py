"""Independent solves on concurrent HIP streams.

Every entry point of the library runs on torch's current stream and waits on
that stream only.  A Lanczos step is a short chain of dependent kernels, a few
microseconds each on a sector of ~10^6 states, so one solve leaves most of the
device idle between launches; solves of independent models issued from
separate host threads onto separate streams overlap on the device.  The
library's calls release the GIL (ctypes), and torch's current stream is per
thread.
"""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor
from typing import Callable, Iterable, List, Optional, TypeVar

import torch

from . import _lib

T = TypeVar("T")
R = TypeVar("R")

__all__ = ["map_streams"]


def map_streams(fn: Callable[[T], R], items: Iterable[T], num_streams: Optional[int] = None) -> List[R]:
    """``[fn(x) for x in items]``, with the calls spread over ``num_streams``
    host threads (default: one per item), each issuing its work on a stream of
    its own (default priority).  Results come back in the order of ``items``;
    every stream is synchronised before the call returns, so device tensors
    in the results are ready on any stream.

    The calls must be independent: distinct models (a model's operator cache
    is not shared between threads), no shared output tensors.  An exception
    in any call is raised here after all threads have finished.
    """
    items = list(items)
    if not items:
        return []
    n = len(items) if num_streams is None else max(1, min(int(num_streams), len(items)))
    _lib.lib()  # load the library before the workers start
    # inputs the caller prepared on its own stream are complete before the
    # workers' streams read them
    torch.cuda.current_stream().synchronize()
    streams = [torch.cuda.Stream() for _ in range(n)]
    results: List[Optional[R]] = [None] * len(items)

    def work(k: int) -> None:
        st = streams[k]
        try:
            with torch.cuda.stream(st):
                for i in range(k, len(items), n):
                    results[i] = fn(items[i])
        finally:
            st.synchronize()

    with ThreadPoolExecutor(max_workers=n) as ex:
        futures = [ex.submit(work, k) for k in range(n)]
        errors = [f.exception() for f in futures]
    for e in errors:
        if e is not None:
            raise e
    return results  # type: ignore[return-value]
