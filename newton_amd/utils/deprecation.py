"""Keyword-only arguments passed positionally: accepted with a ``DeprecationWarning``, as the
reference's ``deprecate_nonkeyword_arguments`` does (``newton/_src/utils/deprecation.py:16-65``)."""

from __future__ import annotations

import functools
import inspect
import warnings


def deprecate_nonkeyword_arguments(func):
    """Wrap ``func`` so that extra positional arguments bind to its keyword-only parameters in order,
    with one ``DeprecationWarning`` naming them; more positionals than there are parameters raise
    ``TypeError`` as a plain call would."""
    sig = inspect.signature(func)
    positional = [p.name for p in sig.parameters.values()
                  if p.kind in (inspect.Parameter.POSITIONAL_ONLY, inspect.Parameter.POSITIONAL_OR_KEYWORD)]
    keyword_only = [p.name for p in sig.parameters.values() if p.kind == inspect.Parameter.KEYWORD_ONLY]
    if not keyword_only:
        return func

    @functools.wraps(func)
    def wrapper(*args, **kwargs):
        extra = args[len(positional):]
        if not extra:
            return func(*args, **kwargs)
        if len(extra) > len(keyword_only):
            sig.bind(*args, **kwargs)  # raises the TypeError of a plain call
        moved = dict(zip(keyword_only, extra))
        clash = sorted(set(moved) & set(kwargs))
        if clash:
            raise TypeError(f"{func.__qualname__}() got multiple values for argument '{clash[0]}'")
        names = ", ".join(f"'{n}'" for n in moved)
        warnings.warn(f"Passing {names} positionally to {func.__qualname__}() is deprecated. "
                      "Pass these arguments as keyword arguments instead.", DeprecationWarning, stacklevel=2)
        return func(*args[:len(positional)], **moved, **kwargs)

    return wrapper


__all__ = ["deprecate_nonkeyword_arguments"]
