"""Attribute policy shared by the faiss stand-ins (the behaviour of
python_interface/python_package/aoclda/faiss/_index_flat_l2.py:146-165 of the reference).

A faiss object has many tunable fields that this backend ignores. Setting a public field
that the class does not list as supported still stores it but warns that it has no effect;
reading a public name the class does not define raises NotImplementedError, so that a
faiss script fails at the first feature the backend lacks rather than later with a wrong
result. Dunder lookups raise AttributeError, as Python's own protocols (copy, pickle)
expect.
"""
import warnings


class _Shim:
    _SUPPORTED_ATTRS = frozenset()

    def __setattr__(self, name, value):
        if not name.startswith("_") and name not in self._SUPPORTED_ATTRS:
            warnings.warn(f"{type(self).__name__}: setting '{name}' has no effect; the "
                          "attribute is not supported by the AOCL-DA MI355X backend.",
                          UserWarning, stacklevel=2)
        object.__setattr__(self, name, value)

    def __getattr__(self, name):
        # Called only when normal lookup fails.
        if name.startswith("__"):
            raise AttributeError(name)
        raise NotImplementedError(
            f"{type(self).__name__}.{name} is not supported by the AOCL-DA MI355X backend")
