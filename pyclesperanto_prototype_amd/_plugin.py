"""The plugin decorator and image typing of the cle API.

Reference: pyclesperanto_prototype/_tier0/_plugin_function.py:12-76 and
_tier0/_types.py:13-23.  Arguments are bound to the signature (TypeError on a
mismatch, as inspect.Signature.bind raises), every argument annotated ``Image`` that
holds an image is pushed to the device as float32, and every ``Image`` argument left
as None is created by ``output_creator`` from the leading bound arguments.
"""
from __future__ import annotations

import inspect
from functools import wraps
from typing import Callable, Union

import numpy as np

from ._array import Array, create_like, push

Image = Union[np.ndarray, Array]


def is_image(obj) -> bool:
    """Reference: pyclesperanto_prototype/_tier0/_types.py:13-23."""
    return isinstance(obj, (np.ndarray, tuple, list, Array)) or str(type(obj)) in (
        "<class 'torch.Tensor'>",
        "<class 'cupy._core.core.ndarray'>",
        "<class 'dask.array.core.Array'>",
        "<class 'xarray.core.dataarray.DataArray'>",
    )


def plugin_function(function: Callable = None, output_creator: Callable = create_like, categories: list = None,
                    priority: int = 0):
    """Decorator; usable bare (@plugin_function) or with keywords, like the reference's curried form."""
    if function is None:
        return lambda f: plugin_function(f, output_creator=output_creator, categories=categories, priority=priority)

    function.fullargspec = inspect.getfullargspec(function)
    function.categories = categories
    function.priority = priority
    sig = inspect.signature(function)
    n_creator = len(inspect.signature(output_creator).parameters)

    @wraps(function)
    def worker_function(*args, **kwargs):
        bound = sig.bind(*args, **kwargs)
        bound.apply_defaults()
        for key, value in bound.arguments.items():
            if is_image(value) and key in sig.parameters and sig.parameters[key].annotation is Image:
                bound.arguments[key] = push(value)
            if key in sig.parameters and sig.parameters[key].annotation is Image and value is None:
                bound.arguments[key] = output_creator(*bound.args[:n_creator])
        return function(*bound.args, **bound.kwargs)

    worker_function.__module__ = "pyclesperanto_prototype_amd"
    return worker_function
