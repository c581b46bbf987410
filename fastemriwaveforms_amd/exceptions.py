"""Exception tree mirroring src/few/utils/exceptions.py:9 and src/few/cutils/__init__.py:9-59."""


class FewException(Exception):
    """Base class of the package's exceptions (utils/exceptions.py:9)."""


class BackendUnavailableException(FewException):
    """The backend is not available (cutils/__init__.py:9)."""


class BackendNotInstalled(BackendUnavailableException):
    """The backend's native library is not built / not loadable (cutils/__init__.py:13)."""


class MissingHardware(BackendUnavailableException):
    """The backend needs a GPU that is not visible (cutils/__init__.py:57)."""


class BackendAccessException(FewException):
    """A requested backend is disabled or unavailable (cutils/__init__.py:547)."""


class FewRocmError(FewException):
    """A native entry point returned an error code (the reference exits the process instead)."""
