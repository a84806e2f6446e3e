"""String-keyed plugin registry (jaxdem/factory.py:28-503).

Each root class (Collider, ForceModel, Domain, ...) owns its registry.  Keys
are normalised at registration and lookup by removing spaces, underscores and
hyphens and lower-casing (factory.py:28-35), so ``"CellList"``,
``"cell_list"`` and ``"Cell-List"`` name the same class.  ``create`` prefers a
``Create`` classmethod over the constructor (factory.py:420-430), raises
``KeyError`` for unknown keys (factory.py:412-418), drops unknown keyword
arguments with a warning (factory.py:485-493) and raises ``TypeError`` for a
bad signature (factory.py:495-501).
"""

from __future__ import annotations

import warnings
from inspect import signature
from typing import Any, Callable, ClassVar, TypeVar

T = TypeVar("T", bound="Factory")


def _normalize_key(key: str) -> str:
    """factory.py:28-35."""
    return key.replace(" ", "").replace("_", "").replace("-", "").lower()


class Factory:
    """Base of every pluggable component."""

    _registry: ClassVar[dict] = {}
    __registry_name__: ClassVar[str | None] = None

    def __init_subclass__(cls, **kw: Any) -> None:
        """factory.py:79-91: every subclass gets its own private registry and
        may not override ``create``."""
        super().__init_subclass__(**kw)
        cls._registry = {}
        cls.__registry_name__ = None
        if "create" in cls.__dict__:
            raise TypeError(
                f"{cls.__name__} is not allowed to override the `create` method. "
                "Use `Create` instead for custom instantiation logic."
            )

    @classmethod
    def registry_name(cls) -> str:
        name = getattr(cls, "__registry_name__", None)
        if name is None:
            raise KeyError(f"{cls.__name__} is not registered in the Factory.")
        return str(name)

    @property
    def type_name(self) -> str:
        return type(self).registry_name()

    @classmethod
    def register(cls, key: str | None = None) -> Callable[[type], type]:
        """factory.py:242-313."""

        def deco(sub: type) -> type:
            k = _normalize_key(sub.__name__ if key is None else key)
            existing = cls._registry.get(k)
            if existing is not None and not (
                existing.__qualname__ == sub.__qualname__ and existing.__module__ == sub.__module__
            ):
                raise ValueError(f"{cls.__name__}: key '{k}' already registered for {existing.__name__}")
            cls._registry[k] = sub
            own = sub.__dict__.get("__registry_name__", None)
            if own is not None and own != k:
                raise ValueError(f"{sub.__name__} has __registry_name__={own!r}, but is being registered as {k!r}.")
            sub.__registry_name__ = k
            return sub

        return deco

    @classmethod
    def create(cls: type[T], key: str, /, **kw: Any) -> T:
        """factory.py:373-503."""
        try:
            sub = cls._registry[_normalize_key(key)]
        except KeyError as err:
            raise KeyError(f"Unknown {cls.__name__} '{key}'. Available: {list(cls._registry)}") from err
        fn = getattr(sub, "Create", None)
        target = fn if fn is not None else sub
        sig = signature(target)
        params = [p for p in sig.parameters if p != "self"]
        var_kw = any(p.kind == p.VAR_KEYWORD for p in sig.parameters.values())
        kw = dict(kw)
        if not var_kw:
            dropped = [p for p in kw if p not in params]
            for p in dropped:
                kw.pop(p)
            if dropped:
                warnings.warn(
                    f"{cls.__name__}.create('{key}'): ignoring unknown keyword(s) {dropped}. "
                    f"Expected signature: {sub.__name__}{sig}",
                    stacklevel=2,
                )
        try:
            sig.bind_partial(**kw)
        except TypeError as err:
            raise TypeError(
                f"Invalid keyword(s) for {sub.__name__}: {err}. Expected signature: {sub.__name__}{sig}"
            ) from None
        return target(**kw)


__all__ = ["Factory", "_normalize_key"]
