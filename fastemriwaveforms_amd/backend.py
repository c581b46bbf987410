"""Backend plugin contract for the rocm backend.

Mirrors the reference's backend system (src/few/cutils/__init__.py:62-670 and
src/few/utils/baseclasses.py:29-119) for the two methods on the interpolated
mode-summation path:

* :class:`BackendMethods` / :class:`Backend` with ``Feature`` flags
  (cutils/__init__.py:62-158),
* :class:`RocmBackend` -- the new backend.  Its ``interpolate_arrays_wrap``
  and ``get_waveform_wrap`` reach the HIP kernels of libfew_rocm.so; it
  declares ``Feature.NUMPY`` with ``xp = numpy`` so FEW's own, unchanged
  ``CubicSplineInterpolant`` / ``InterpolatedModeSum`` drive it with host
  arrays (see :mod:`fastemriwaveforms_amd.pyinterp`).  The other FEW
  backend methods (AAK, amplitude interpolation, ROMAN layers, FD sum) raise.
* :data:`KNOWN_BACKENDS`, :class:`BackendsManager`, :func:`get_backend`,
  :func:`get_first_backend` -- registry semantics of cutils/__init__.py:514-667
  (enable / disable, first-available selection, errors).
* :class:`ParallelModuleBase` -- backend selection of baseclasses.py:29-119,
  with ``"rocm"`` first in the GPU preference lists.
"""

from __future__ import annotations

import dataclasses
import enum
import types
import typing

from .exceptions import (BackendAccessException, BackendNotInstalled,
                         BackendUnavailableException, MissingHardware)


@dataclasses.dataclass
class BackendMethods:
    """cutils/__init__.py:62-71"""

    pyWaveform: typing.Callable
    interp2D: typing.Callable
    interpolate_arrays_wrap: typing.Callable
    get_waveform_wrap: typing.Callable
    get_waveform_generic_fd_wrap: typing.Callable
    neural_layer_wrap: typing.Callable
    transform_output_wrap: typing.Callable
    xp: types.ModuleType


class Backend:
    """Abstract backend (cutils/__init__.py:74-158)."""

    class Feature(enum.Flag):
        NUMPY = enum.auto()
        CUPY = enum.auto()
        CUDA = enum.auto()
        GPU = enum.auto()
        NONE = 0

    def __init__(self, name: str, methods: BackendMethods, features: "Backend.Feature"):
        self.name = name
        self.pyWaveform = methods.pyWaveform
        self.interp2D = methods.interp2D
        self.interpolate_arrays_wrap = methods.interpolate_arrays_wrap
        self.get_waveform_wrap = methods.get_waveform_wrap
        self.get_waveform_generic_fd_wrap = methods.get_waveform_generic_fd_wrap
        self.neural_layer_wrap = methods.neural_layer_wrap
        self.transform_output_wrap = methods.transform_output_wrap
        self.xp = methods.xp
        self.features = features

    def supports(self, feature: "Backend.Feature") -> bool:
        return feature in self.features

    @property
    def uses_gpu(self) -> bool:
        return self.supports(Backend.Feature.GPU)

    @property
    def uses_numpy(self) -> bool:
        return self.supports(Backend.Feature.NUMPY)

    @property
    def uses_cupy(self) -> bool:
        return self.supports(Backend.Feature.CUPY)

    @property
    def uses_cuda(self) -> bool:
        return self.supports(Backend.Feature.CUDA)


def _not_in_backend(name: str):
    def fn(*args, **kwargs):
        raise NotImplementedError(f"{name} is not provided by the rocm backend")
    fn.__name__ = name
    return fn


class RocmBackend(Backend):
    """The MI355X backend: HIP kernels behind libfew_rocm.so."""

    @staticmethod
    def rocm_methods_loader() -> BackendMethods:
        try:
            import numpy
            from . import _lib, pyinterp
        except (ModuleNotFoundError, ImportError) as e:
            raise BackendUnavailableException("'rocm' backend could not be imported.") from e
        _lib.lib()  # raises BackendNotInstalled when libfew_rocm.so is missing
        return BackendMethods(
            pyWaveform=pyinterp.pyWaveform,
            interp2D=pyinterp.interp2D,
            interpolate_arrays_wrap=pyinterp.interpolate_arrays_wrap,
            get_waveform_wrap=pyinterp.get_waveform_wrap,
            get_waveform_generic_fd_wrap=pyinterp.get_waveform_generic_fd_wrap,
            neural_layer_wrap=_not_in_backend("neural_layer_wrap"),
            transform_output_wrap=_not_in_backend("transform_output_wrap"),
            xp=numpy,
        )

    def __init__(self):
        methods = self.rocm_methods_loader()
        import torch
        if not torch.cuda.is_available():
            raise MissingHardware("'rocm' backend needs a visible AMD GPU.")
        super().__init__(name="rocm", methods=methods, features=Backend.Feature.NUMPY)

    @property
    def precision(self) -> str:
        """Arithmetic of get_waveform_wrap ("fp64", "fact64", "mixed16", "mixed"); the
        FEW_ROCM_PRECISION environment variable sets it at import."""
        from . import pyinterp
        return pyinterp.get_precision()

    @precision.setter
    def precision(self, value: str) -> None:
        from . import pyinterp
        pyinterp.set_precision(value)


KNOWN_BACKENDS = {"rocm": RocmBackend}


class BackendStatus:
    pass


class BackendStatusUnloaded(BackendStatus):
    pass


@dataclasses.dataclass
class BackendStatusLoaded(BackendStatus):
    instance: Backend


class BackendStatusDisabled(BackendStatus):
    pass


@dataclasses.dataclass
class BackendStatusUnavailable(BackendStatus):
    reason: BackendUnavailableException


class BackendsManager:
    """Backend registry (cutils/__init__.py:551-667)."""

    def __init__(self, enabled_backends: typing.Optional[typing.Sequence[str]] = None,
                 known: typing.Optional[dict] = None):
        self._known = dict(KNOWN_BACKENDS if known is None else known)
        self._registry = {name: BackendStatusUnloaded() for name in self._known}
        if enabled_backends is not None:
            for name in self.backend_list:
                if name in enabled_backends:
                    status = self._try_loading_backend(name)
                    if isinstance(status, BackendStatusUnavailable):
                        raise BackendAccessException(
                            f"Backend '{name}' is marked as enabled but cannot be loaded"
                        ) from status.reason
                else:
                    self._registry[name] = BackendStatusDisabled()

    @property
    def backend_list(self) -> typing.List[str]:
        return list(self._registry.keys())

    def _try_loading_backend(self, name: str) -> BackendStatus:
        if name not in self._registry:
            raise ValueError("'{}' is not a valid backend name, expected any of: {}".format(
                name, ", ".join(self.backend_list)))
        status = self._registry[name]
        if not isinstance(status, BackendStatusUnloaded):
            return status
        try:
            status = BackendStatusLoaded(instance=self._known[name]())
        except BackendUnavailableException as e:
            status = BackendStatusUnavailable(reason=e)
        self._registry[name] = status
        return status

    def get_backend(self, name: str) -> Backend:
        status = self._try_loading_backend(name)
        if isinstance(status, BackendStatusDisabled):
            raise BackendAccessException(
                f"Backend '{name}' cannot be accessed, it has been disabled.")
        if isinstance(status, BackendStatusUnavailable):
            raise BackendAccessException(
                f"Backend '{name}' is unavailable. See previous error messages.") from status.reason
        return status.instance

    def has_backend(self, name: str) -> bool:
        return isinstance(self._try_loading_backend(name), BackendStatusLoaded)

    def get_first_backend(self, backends: typing.Sequence[str]) -> Backend:
        assert len(backends) > 0
        reasons = []
        for name in backends:
            if name not in self._registry:
                continue
            status = self._try_loading_backend(name)
            if isinstance(status, BackendStatusLoaded):
                return status.instance
            if isinstance(status, BackendStatusUnavailable):
                reasons.append(status.reason)
            if isinstance(status, BackendStatusDisabled):
                reasons.append(BackendAccessException(f"Backend '{name}' is disabled."))
        raise BackendAccessException(
            "Could not access any of the following backends which are either disabled or "
            "unavailable: {}".format(", ".join(backends))) from (reasons[0] if reasons else None)


_MANAGER: typing.Optional[BackendsManager] = None


def backends_manager() -> BackendsManager:
    global _MANAGER
    if _MANAGER is None:
        _MANAGER = BackendsManager()
    return _MANAGER


def get_backend(name: str) -> Backend:
    """utils/globals.py:378-389"""
    return backends_manager().get_backend(name)


def get_first_backend(names: typing.Sequence[str]) -> Backend:
    """utils/globals.py:406-408"""
    return backends_manager().get_first_backend(names)


class ParallelModuleBase:
    """Backend selection (utils/baseclasses.py:29-119) with "rocm" preferred."""

    def __init__(self, force_backend=None):
        if force_backend is not None:
            if isinstance(force_backend, Backend):
                force_backend = force_backend.name
            self._backend_name = get_backend(force_backend).name
        else:
            self._backend_name = get_first_backend(self.supported_backends()).name

    @property
    def backend(self) -> Backend:
        return get_backend(self._backend_name)

    @classmethod
    def supported_backends(cls):
        raise NotImplementedError(
            "Class {} does not implement the supported_backends method.".format(cls))

    @staticmethod
    def GPU_RECOMMENDED():
        return ["rocm"]

    @staticmethod
    def GPU_ONLY():
        return ["rocm"]

    @property
    def xp(self):
        return self.backend.xp

    @property
    def backend_name(self) -> str:
        return self.backend.name

    def build_with_same_backend(self, module_class, args=None, kwargs=None):
        args = [] if args is None else args
        kwargs = {} if kwargs is None else dict(kwargs)
        kwargs["force_backend"] = self.backend_name
        return module_class(*args, **kwargs)
