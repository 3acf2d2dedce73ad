"""The reference's nipype interface (python/mirorr/interface.py:10-161),
restated without nipype (not installed here) so that its command lines and
output lists can be produced and run against this build's `mirorr`
executable.

Mirorr().inputs carries the same input names, defaults and argument strings;
cmdline is assembled the way nipype's CommandLine does it: inputs in
alphabetical order, a bool input emits its flag when true, other inputs
format their argument string with their value, and `transform` is generated
as <transform_type>.tfm in the working directory when not given.
run() executes the command; _list_outputs() returns nipype's output dict.
"""
from __future__ import annotations

import os
import shlex
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_CMD = os.path.join(_HERE, "bin", "mirorr")


class _Undefined:
    def __repr__(self):
        return "<undefined>"


Undefined = _Undefined()


def isdefined(v) -> bool:
    return v is not Undefined


# name: (argstr, kind, default-or-Undefined, usedefault, extra)
#   kind: "file" (must exist), "outfile", "bool", "int", "enum"
_SPEC = {
    "moving_image": ("--moving %s", "file", Undefined, False, None),
    "moving_image_mask": ("--moving-mask %s", "file", Undefined, False, None),
    "fixed_image": ("--fixed %s", "file", Undefined, False, None),
    "fixed_image_mask": ("--fixed-mask %s", "file", Undefined, False, None),
    "initial_moving_transform": ("--start-tfm %s", "file", Undefined, False, None),
    "invert_initial_moving_transform": ("--invert-start-tfm", "bool", Undefined, False, None),
    "transform": ("--last-tfm %s", "outfile", Undefined, False, None),
    "invert_output_transform": ("--invert-last-tfm", "bool", Undefined, False, None),
    "transform_type": ("--tfm-type %s", "enum", "rigid", True, ("rigid", "affine")),
    "perform_registration": ("--do-not-register", "bool", True, False, None),
    "reoriented_fixed": ("--save-reoriented-fixed %s", "outfile", Undefined, False,
                         "resampled_fixed"),
    "reoriented_moving": ("--save-reoriented-moving %s", "outfile", Undefined, False,
                          "resampled_moving"),
    "resampled_fixed": ("--save-resampled-fixed %s", "outfile", Undefined, False,
                        "reoriented_fixed"),
    "resampled_moving": ("--save-resampled-moving %s", "outfile", Undefined, False,
                         "reoriented_moving"),
    "interpolation": ("--final-interpolater %s", "enum", Undefined, False,
                      ("Linear", "NearestNeighbor", "Sinc", "BSpline")),
    "pyramid_start": ("--pyr-start %d", "int", 1, False, None),
    "pyramid_end": ("--pyr-end %d", "int", Undefined, False, None),
    "number_of_pyramid_levels": ("--pyr-num", "int", 1024, False, None),
    "number_of_iterations": ("--iterations %d", "int", Undefined, False, None),
    "metric": ("--blockmetric %s", "enum", Undefined, False, ("NC", "SD", "CR", "MI")),
    "block_width": ("--blockwidth %d", "int", Undefined, False, None),
}
_MANDATORY = ("moving_image", "fixed_image")


class MirorrInputs:
    """Input spec of MirorrInputSpec (python/mirorr/interface.py:10-87)."""

    def __init__(self):
        object.__setattr__(self, "_set", {})

    def __getattr__(self, name):
        if name not in _SPEC:
            raise AttributeError(name)
        return self._set.get(name, Undefined)

    def __setattr__(self, name, value):
        if name not in _SPEC:
            raise AttributeError(f"Mirorr has no input '{name}'")
        argstr, kind, _, _, extra = _SPEC[name]
        if value is Undefined:
            self._set.pop(name, None)
            return
        if kind == "file" and not os.path.exists(value):
            raise FileNotFoundError(f"the '{name}' input must be an existing file: {value}")
        if kind == "bool" and not isinstance(value, bool):
            raise TypeError(f"the '{name}' input must be a bool")
        if kind == "int" and (not isinstance(value, int) or isinstance(value, bool)):
            raise TypeError(f"the '{name}' input must be an int")
        if kind == "enum" and value not in extra:
            raise ValueError(f"the '{name}' input must be one of {extra}")
        self._set[name] = value

    def value(self, name):
        """The value nipype formats: the set one, or the default when the
        input uses it (transform_type only)."""
        if name in self._set:
            return self._set[name]
        _, _, default, usedefault, _ = _SPEC[name]
        return default if usedefault else Undefined


class Mirorr:
    """Mirror-symmetric block-matching registration (nipype CommandLine
    interface of python/mirorr/interface.py:100-161)."""

    _cmd = DEFAULT_CMD

    def __init__(self, cwd: str | None = None, **inputs):
        self.inputs = MirorrInputs()
        self.cwd = cwd or os.getcwd()
        for k, v in inputs.items():
            setattr(self.inputs, k, v)

    # -- nipype hooks of the reference ------------------------------------
    def _gen_filename(self, name):
        if name == "transform":
            t = self.inputs.value("transform")
            if isdefined(t):
                return t
            return os.path.join(self.cwd, "{}.tfm".format(self.inputs.value("transform_type")))
        return None

    def _format_arg(self, name, value):
        argstr, kind, _, _, _ = _SPEC[name]
        if name == "perform_registration":
            value = not value  # --do-not-register when registration is off
        if name == "interpolation":
            value = dict(Linear="linear", NearestNeighbor="nn", BSpline="bspline",
                         Sinc="sinc")[value]
        if name == "metric":
            value = value.lower()
        if kind == "bool":
            return argstr if value else None
        # "--pyr-num" has no format field in the reference's spec: nipype's
        # `argstr % value` raises there.
        return argstr % value

    def _check(self):
        for name in _MANDATORY:
            if not isdefined(self.inputs.value(name)):
                raise ValueError(f"Mirorr requires a value for input '{name}'")
        for name, (_, _, _, _, extra) in _SPEC.items():
            if (isinstance(extra, str) and isdefined(self.inputs.value(name))
                    and isdefined(self.inputs.value(extra))):
                raise ValueError(f"inputs '{name}' and '{extra}' are mutually exclusive")

    @property
    def cmdline(self) -> str:
        self._check()
        args = []
        for name in sorted(_SPEC):
            value = self.inputs.value(name)
            if not isdefined(value):
                if _SPEC[name][1] == "outfile" and name == "transform":
                    value = self._gen_filename(name)
                else:
                    continue
            s = self._format_arg(name, value)
            if s:
                args.append(s)
        return " ".join([shlex.quote(self._cmd), *args])

    def _list_outputs(self):
        outputs = {}
        outputs["reverse_transforms"] = outputs["forward_transforms"] = [
            self._gen_filename("transform")]
        outputs["forward_invert_flags"] = [False]
        outputs["reverse_invert_flags"] = [True]
        for output_name in ("reoriented_fixed", "reoriented_moving", "resampled_fixed",
                            "resampled_moving"):
            v = self.inputs.value(output_name)
            if isdefined(v):
                outputs[output_name] = os.path.abspath(os.path.join(self.cwd, v))
        # The reference's aliases, as written there: each one reads the other
        # output's key (a KeyError when only one of the two is set).
        if "resampled_fixed" in outputs:
            outputs["warped_image"] = outputs["resampled_moving"]
        if "resampled_moving" in outputs:
            outputs["inverse_warped_image"] = outputs["resampled_fixed"]
        return outputs

    def run(self, timeout: float = 600.0):
        """Runs the command line in self.cwd; returns (returncode, stdout,
        stderr, outputs)."""
        r = subprocess.run(shlex.split(self.cmdline), cwd=self.cwd, capture_output=True,
                           text=True, timeout=timeout)
        outputs = self._list_outputs() if r.returncode == 0 else {}
        return r.returncode, r.stdout, r.stderr, outputs
