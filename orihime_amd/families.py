"""The alignment families as data, and everything public built from them.

Each family (local Smith-Waterman, global Needleman-Wunsch, Saigo-Vert; each
with a linear or an affine gap model) is one ``Family`` record: its name,
its gap arguments, its kernel adapters in ``ops/`` and a few words for the
documentation.  From the records this module builds, with one code path for
all families:

* the public functions ``<name>``, ``<name>_value``, ``<name>_entropy``
  (reference surface: orihime/sw.py, orihime/nw.py, orihime/sv.py);
* the named operations ``ops.<name>`` with ``forward`` / ``backward`` /
  ``sensitivity`` and explicit output selection (orihime/_public_ops.py);
* the ``nn`` layers with per-parameter learnability (orihime/nn.py);
* the ``raw`` tier's parameter VJPs (orihime/raw/*.py).

Public signatures, argument names and defaults, and the error messages that
callers match on are the reference's; the structure and the text are this
package's.
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass
from importlib import import_module
from inspect import Parameter, Signature
from types import MappingProxyType
from typing import Any

import torch
from torch import Tensor

from . import _checks
from .operator import Operator

# --------------------------------------------------------------------------
# Records
# --------------------------------------------------------------------------

#: public gap argument -> its documentation
_GAP_ARGS = {
    "gap_score": "score of every gap position (insertion or deletion)",
    "gap_open_score": "score of the first position of each gap",
    "gap_extend_score": "score of each further position of a gap",
}
#: public parameter name -> the engine's (and the kernel schemas') name
_ENGINE_NAME = {"gap_score": "gap", "gap_open_score": "gap_open",
                "gap_extend_score": "gap_ext", "temperature": "temp"}


@dataclass(frozen=True)
class Family:
    """One alignment family."""

    name: str            # public prefix and ops name
    adapter: str         # module under ops/ holding the kernel adapters
    kernels: str         # OperatorKernels attribute of that module
    gaps: tuple[str, ...]
    mode: str            # how alignments may start and end (documentation)
    title: str           # human name (documentation)
    layer: str           # nn class name
    layer_alias: str

    @property
    def affine(self) -> bool:
        return len(self.gaps) == 2

    @property
    def parameters(self) -> tuple[str, ...]:
        """Public scalar parameters, temperature last."""
        return (*self.gaps, "temperature")

    @property
    def engine_parameters(self) -> tuple[str, ...]:
        return tuple(_ENGINE_NAME[p] for p in self.parameters)


_LOCAL = ("local: an alignment may start and end anywhere in either sequence; the "
          "empty alignment counts too")
_GLOBAL = "global: every alignment covers both sequences from their first to their last position"
_SV = ("local in the Saigo-Vert form: alignments start and end with a match, and the empty "
       "alignment counts once")

FAMILIES: dict[str, Family] = {f.name: f for f in (
    Family("sw", "sw", "kernels", ("gap_score",), _LOCAL, "Smith-Waterman",
           "SmithWaterman", ""),
    Family("sw_affine", "sw_affine", "SW_AFFINE_KERNELS", ("gap_open_score", "gap_extend_score"),
           _LOCAL, "affine-gap Smith-Waterman", "SmithWatermanAffine", ""),
    Family("nw", "nw", "NW_KERNELS", ("gap_score",), _GLOBAL, "Needleman-Wunsch",
           "NeedlemanWunsch", "NW"),
    Family("nw_affine", "nw_affine", "NW_AFFINE_KERNELS", ("gap_open_score", "gap_extend_score"),
           _GLOBAL, "affine-gap Needleman-Wunsch", "NeedlemanWunschAffine", "NWAffine"),
    Family("sv", "sv_linear", "SV_KERNELS", ("gap_score",), _SV, "linear-gap Saigo-Vert",
           "SaigoVertLinear", "SVLinear"),
    Family("sv_affine", "sv_affine", "SV_AFFINE_KERNELS", ("gap_open_score", "gap_extend_score"),
           _SV, "affine-gap Saigo-Vert", "SaigoVertAffine", "SVAffine"),
)}

_ENGINES: dict[str, Operator] = {}


def engine(name: str) -> Operator:
    """The autograd engine (``operator.Operator``) of family ``name``."""
    op = _ENGINES.get(name)
    if op is None:
        fam = FAMILIES[name]
        adapters = import_module(f".ops.{fam.adapter}", __package__)
        op = Operator(fam.name, params=fam.engine_parameters, kernels=getattr(adapters, fam.kernels))
        _ENGINES[name] = op
    return op


# --------------------------------------------------------------------------
# Public functions
# --------------------------------------------------------------------------

_KW = Parameter.KEYWORD_ONLY


def signature_of(fam: Family) -> Signature:
    """``(pair_scores, *, <gaps>=0.0, temperature=1.0, lengths=None,
    mask=None, dtype=None)``."""
    params = [Parameter("pair_scores", Parameter.POSITIONAL_OR_KEYWORD, annotation=Tensor)]
    params += [Parameter(g, _KW, default=0.0, annotation=float | Tensor) for g in fam.gaps]
    params += [Parameter("temperature", _KW, default=1.0, annotation=float | Tensor),
               Parameter("lengths", _KW, default=None, annotation=Tensor | None),
               Parameter("mask", _KW, default=None, annotation=Tensor | None),
               Parameter("dtype", _KW, default=None, annotation=torch.dtype | None)]
    return Signature(params, return_annotation=Tensor)


_ARG_TEXT = {
    "pair_scores": ("FP32 tensor ``[B, L1, L2]``: the score of aligning position i of the first "
                    "sequence with position j of the second (higher favours the match); "
                    "``-inf`` removes a cell."),
    "temperature": ("softness of the maximum over alignments; finite and > 0 (a number or a "
                    "one-element FP32 tensor).  Scores and gaps must satisfy "
                    "``abs(x) / temperature <= 80``."),
    "lengths": ("optional contiguous ``torch.int32`` tensor ``[B, 2]`` on the scores' device: "
                "the rows and columns of each pair that take part; the rest of the padded "
                "input is ignored (its map entries are 0)."),
    "mask": ("optional bool tensor shaped like ``pair_scores``; ``True`` takes that cell out "
             "of every alignment."),
    "dtype": ("optional compute dtype; only ``torch.float32`` is accepted (lower-precision "
              "floating inputs are cast to it)."),
}

_WHAT = {
    "map": ("Posterior match probabilities of soft {title} alignment: entry (b, i, j) is "
            "the probability, under the Gibbs distribution over alignments at the given "
            "temperature, that pair b aligns position i with position j.",
            "FP32 tensor shaped like ``pair_scores``."),
    "value": ("Soft {title} score: temperature times the log of the sum of exp(score / "
              "temperature) over all alignments, a smooth maximum whose gradient with respect "
              "to ``pair_scores`` is the match-probability map.",
              "FP32 tensor ``[B]``, one value per pair."),
    "entropy": ("Entropy of the Gibbs distribution over soft {title} alignments (the "
                "derivative of the value with respect to the temperature).  Gradients flow to "
                "``pair_scores`` only.",
                "FP32 tensor ``[B]``, one value per pair."),
}


def _doc(fam: Family, kind: str) -> str:
    summary, returns = _WHAT[kind]
    lines = [summary.format(title=fam.title), "", f"Alignments are {fam.mode}.", "", "Args:"]
    for name in signature_of(fam).parameters:
        text = _GAP_ARGS[name] + " (a number or a one-element FP32 tensor)." \
            if name in _GAP_ARGS else _ARG_TEXT[name]
        lines.append(f"    {name}: {text}")
    lines += ["", "Returns:", f"    {returns}", "", "Raises:",
              "    TypeError: an input of the wrong kind or dtype.",
              "    ValueError: a shape, device, length or mask that does not fit the scores.",
              "    RuntimeError: a score, gap or temperature outside the numerical domain, or "
              "lengths beyond the padded shape."]
    return "\n".join(lines)


def _public(fam: Family, kind: str):
    sig = signature_of(fam)
    eng = engine(fam.name)
    run = {"map": eng.__call__, "value": eng.value, "entropy": eng.entropy}[kind]
    names = fam.parameters

    def function(*args: Any, **kwargs: Any) -> Tensor:
        a = sig.bind(*args, **kwargs)
        a.apply_defaults()
        v = a.arguments
        return run(v["pair_scores"], *(v[n] for n in names), lengths=v["lengths"],
                   mask=v["mask"], dtype=v["dtype"])

    function.__name__ = function.__qualname__ = fam.name + ("" if kind == "map" else f"_{kind}")
    function.__module__ = __package__
    function.__signature__ = sig
    function.__doc__ = _doc(fam, kind)
    return function


_PUBLIC: dict[str, tuple] = {}


def public_functions(name: str) -> tuple:
    """(map, value, entropy) public functions of family ``name``."""
    if name not in _PUBLIC:
        _PUBLIC[name] = tuple(_public(FAMILIES[name], k) for k in ("map", "value", "entropy"))
    return _PUBLIC[name]


# --------------------------------------------------------------------------
# Named operations (ops.<name>)
# --------------------------------------------------------------------------


def _selection(output, allowed: tuple[str, ...]) -> tuple[tuple[str, ...], bool]:
    """The fields ``output`` selects and whether it named a single one."""
    if output is None:
        return allowed, False
    if isinstance(output, str):
        chosen, single = (output,), True
    elif isinstance(output, Sequence):
        chosen, single = tuple(output), False
    else:
        raise TypeError("output must be None, a field name, or a sequence of field names")
    if len(chosen) == 0:
        raise ValueError("output must select at least one field")
    if any(not isinstance(c, str) for c in chosen):
        raise TypeError("output field names must be strings")
    if len(set(chosen)) < len(chosen):
        raise ValueError(f"output must not contain duplicate field names, got {chosen!r}")
    for c in chosen:
        if c not in allowed:
            raise ValueError(f"unknown output field {c!r}; expected one of {allowed!r}")
    return chosen, single


def _with(sig: Signature, *extra: Parameter) -> Signature:
    return sig.replace(parameters=[*sig.parameters.values(), *extra])


class NamedOperation:
    """``forward`` (map / value / entropy), ``backward`` (the map's VJP for
    the scores and each scalar parameter) and ``sensitivity`` (the map's
    derivative field for each scalar parameter) of one family, each
    returning a dict of the selected fields or, for ``output="<field>"``,
    that one tensor."""

    outputs = ("map", "value", "entropy")

    def __init__(self, fam: Family) -> None:
        self.name = fam.name
        self._family = fam
        self._sig = signature_of(fam)
        self.input_fields = ("pair_scores",)
        self.parameter_fields = fam.parameters
        self.backward_fields = (*self.input_fields, *self.parameter_fields)
        self.sensitivity_fields = self.parameter_fields
        out = Parameter("output", _KW, default=None)
        for method, sig in (("forward", _with(self._sig, out)),
                            ("backward", _with(self._sig, Parameter("grad_map", _KW), out)),
                            ("sensitivity", _with(self._sig, out))):
            bound = getattr(self, "_" + method)

            def call(*args, _m=bound, **kwargs):
                return _m(*args, **kwargs)

            call.__name__ = method
            call.__qualname__ = f"ops.{fam.name}.{method}"
            call.__signature__ = sig
            call.__doc__ = bound.__doc__
            setattr(self, method, call)

    def _bind(self, args, kwargs) -> dict[str, Any]:
        a = self._sig.bind(*args, **kwargs)
        a.apply_defaults()
        v = dict(a.arguments)
        if not isinstance(v["pair_scores"], Tensor):
            raise TypeError("pair_scores must be a tensor")
        return v

    def _forward(self, *args, output=None, **kwargs):
        """Selected outputs among ``map``, ``value`` and ``entropy``."""
        chosen, single = _selection(output, self.outputs)
        fns = dict(zip(self.outputs, public_functions(self.name)))
        res = {c: fns[c](*args, **kwargs) for c in chosen}
        return res[chosen[0]] if single else res

    def _backward(self, *args, grad_map: Tensor, output=None, **kwargs):
        """The map's cotangent ``grad_map`` pulled back to the selected
        inputs: ``pair_scores`` (a tensor like it) and the scalar parameters
        (each shaped like the argument given, or ``[]`` for numbers)."""
        chosen, single = _selection(output, self.backward_fields)
        v = self._bind(args, kwargs)
        eng = engine(self.name)
        params = [v[p] for p in self.parameter_fields]
        res = {}
        for c in chosen:
            wrt = "pair_scores" if c == "pair_scores" else _ENGINE_NAME[c]
            with torch.no_grad():
                g = eng._vjp_one(v["pair_scores"], *params, grad_map=grad_map, wrt=wrt,
                                 lengths=v["lengths"], dtype=v["dtype"], mask=v["mask"])
            if c != "pair_scores":
                given = v[c]
                g = (g.to(device=given.device, dtype=given.dtype).reshape(given.shape)
                     if isinstance(given, Tensor) else g.reshape(()))
            res[c] = g
        return res[chosen[0]] if single else res

    def _sensitivity(self, *args, output=None, **kwargs):
        """d(map)/d(parameter) for the selected scalar parameters, each shaped
        like ``pair_scores``."""
        chosen, single = _selection(output, self.sensitivity_fields)
        v = self._bind(args, kwargs)
        eng = engine(self.name)
        params = [v[p] for p in self.parameter_fields]
        res = {}
        for c in chosen:
            with torch.no_grad():
                res[c] = eng._sensitivity_one(v["pair_scores"], *params, wrt=_ENGINE_NAME[c],
                                              lengths=v["lengths"], dtype=v["dtype"],
                                              mask=v["mask"])
        return res[chosen[0]] if single else res

    def __repr__(self) -> str:
        return f"<orihime_amd.ops.{self.name}>"


def install_operations(namespace) -> None:
    """Give ``namespace`` (the ``ops`` package) one NamedOperation per family
    and, under ``_kernels``, the kernel adapter modules by native name."""
    for name, fam in FAMILIES.items():
        setattr(namespace, name, NamedOperation(fam))
    namespace._kernels = MappingProxyType({
        fam.adapter: import_module(f".ops.{fam.adapter}", __package__) for fam in FAMILIES.values()})
    namespace.__all__ = ["sw", "sw_affine", "nw_affine", "nw", "sv", "sv_affine"]


# --------------------------------------------------------------------------
# Raw tier: parameter VJPs without autograd graphs
# --------------------------------------------------------------------------


def raw_vjp(name: str):
    """(vjp_fields, vjp_one, vjp) for family ``name``: the map's cotangent
    contracted with the map's derivative for the selected scalar
    parameters, computed without building an autograd graph."""
    fam = FAMILIES[name]
    fields = fam.parameters
    sig = Signature([
        Parameter("pair_scores", Parameter.POSITIONAL_OR_KEYWORD),
        Parameter("wrt", _KW), Parameter("cotangent", _KW),
        *(Parameter(p, _KW, default=0.0) for p in fam.gaps),
        Parameter("temperature", _KW, default=1.0), Parameter("lengths", _KW, default=None)])

    def one(v: dict[str, Any], wrt: str) -> Tensor:
        if not isinstance(wrt, str):
            raise TypeError("wrt must be a string")
        if wrt not in fields:
            raise ValueError(f"invalid wrt field {wrt!r}; expected one of {fields!r}")
        with torch.no_grad():
            return engine(name)._vjp_one(v["pair_scores"], *(v[p] for p in fields),
                                         grad_map=v["cotangent"], wrt=_ENGINE_NAME[wrt],
                                         lengths=v["lengths"])

    def bind(args, kwargs) -> dict[str, Any]:
        a = sig.bind(*args, **kwargs)
        a.apply_defaults()
        v = dict(a.arguments)
        _checks.derivative_vector(v["pair_scores"], v["cotangent"], name="cotangent",
                                  primary_name="map", error_type=None)
        return v

    def vjp_one(*args, **kwargs) -> Tensor:
        v = bind(args, kwargs)
        return one(v, v["wrt"])

    def vjp(*args, **kwargs) -> dict[str, Tensor]:
        v = bind(args, kwargs)
        wrt = v["wrt"]
        if not isinstance(wrt, tuple):
            raise TypeError("wrt must be a tuple of field names")
        if len(set(wrt)) < len(wrt):
            raise ValueError("wrt must not contain duplicates")
        for w in wrt:
            if not isinstance(w, str) or w not in fields:
                one(v, w)  # raises the field error
        return {w: one(v, w) for w in wrt}

    for fn, kind in ((vjp_one, "one selected"), (vjp, "the selected")):
        fn.__signature__ = sig
        fn.__doc__ = (f"Return {kind} {fam.title} parameter VJP"
                      f"{'' if fn is vjp_one else 's in a dict'}: "
                      "sum over cells of cotangent * d(map)/d(parameter).")
    vjp_one.__name__, vjp.__name__ = "vjp_one", "vjp"
    return fields, vjp_one, vjp


__all__ = ["FAMILIES", "Family", "NamedOperation", "engine", "install_operations",
           "public_functions", "raw_vjp", "signature_of"]
