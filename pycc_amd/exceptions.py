"""Exception types of the public API, mirroring pycc's (``pycc/exceptions.py:13-50``).

``PyCCError`` is the base error, ``PyCCWarning`` the base warning and ``InvalidKeywordError`` a
keyword whose value is not allowed (also a ``ValueError``). ``DeviceError`` carries a failure
reported by the HIP library: its status code and message.
"""
from __future__ import annotations


class PyCCError(Exception):
    """Base class for all errors raised by this package."""


class PyCCWarning(UserWarning):
    """Base class for all warnings raised by this package."""


class InvalidKeywordError(PyCCError, ValueError):
    """A keyword argument was given a value outside the allowed set."""

    def __init__(self, keyword, value, allowed):
        self.keyword = keyword
        self.value = value
        self.allowed = list(allowed)
        allowed_str = ", ".join(repr(a) for a in self.allowed)
        super().__init__("%r is not an allowed value for '%s'. Allowed values: %s."
                         % (value, keyword, allowed_str))


class DeviceError(PyCCError):
    """A call into the HIP library failed; ``code`` is its negative status."""

    def __init__(self, code, message, call=None):
        self.code = code
        self.call = call
        super().__init__("%s failed (status %d): %s" % (call or "ccx call", code, message))
