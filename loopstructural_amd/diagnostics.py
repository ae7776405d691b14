"""Set-up diagnostics of a discrete interpolator.

The three record types carry the field names of the reference's report
(LI/_diagnostics.py:7-102) so that code reading ``report.families[name]
.row_count`` or ``report.to_dict()`` keeps working; everything else here is
generated from the dataclass fields: ``to_dict`` walks the fields in
declaration order (nested records become dicts), and ``summary`` renders the
report as an aligned text table.
"""
from __future__ import annotations

from dataclasses import dataclass, field, fields, is_dataclass


def _plain(value):
    """Nested records -> dicts, dicts of records -> dicts of dicts."""
    if is_dataclass(value):
        return value.to_dict()
    if isinstance(value, dict):
        return {k: _plain(v) for k, v in value.items()}
    return value


class _Record:
    """Dict form of a frozen dataclass; fields listed in ``_hidden`` stay out
    (a family's name is the key it is stored under)."""

    _hidden: tuple = ()

    def to_dict(self) -> dict:
        return {f.name: _plain(getattr(self, f.name)) for f in fields(self)
                if f.name not in self._hidden}


@dataclass(frozen=True)
class ConstraintFamilyDiagnostics(_Record):
    """Rows one constraint family contributed to the least-squares system."""

    _hidden = ("name",)
    name: str
    active: bool
    row_count: int
    dropped_rows: int | None
    effective_weight_mean: float | None
    effective_weight_min: float | None
    effective_weight_max: float | None
    source_point_count: int = 0
    outside_model_point_count: int = 0


@dataclass(frozen=True)
class RegionCoverageDiagnostics(_Record):
    """How many support nodes the active region holds."""

    total_support_nodes: int
    active_region_nodes: int
    inactive_region_nodes: int
    active_fraction: float


@dataclass(frozen=True)
class ConstraintDiagnosticsReport(_Record):
    """Families by name, region coverage and per-family outside-point counts."""

    interpolator_type: str
    families: dict = field(default_factory=dict)
    region_coverage: RegionCoverageDiagnostics | None = None
    outside_model_points: dict = field(default_factory=dict)

    @property
    def total_rows(self) -> int:
        return int(sum(f.row_count for f in self.families.values()))

    @property
    def active_families(self) -> dict:
        return {k: v for k, v in self.families.items() if v.active}

    def to_dict(self) -> dict:
        d = super().to_dict()
        # the row total is derived, listed right after the type
        return {"interpolator_type": d.pop("interpolator_type"), "total_rows": self.total_rows,
                **d}

    def summary(self) -> str:
        """Text table: one line per active family (rows, dropped rows, weight
        range), then the region coverage and the outside-point counts."""
        head = ("family", "rows", "dropped", "weight mean", "weight range")
        body = []
        for name in sorted(self.active_families):
            f = self.families[name]
            w = [f.effective_weight_mean, f.effective_weight_min, f.effective_weight_max]
            fmt = ["-" if v is None else f"{v:.4g}" for v in w]
            body.append((name, str(f.row_count), "-" if f.dropped_rows is None
                         else str(f.dropped_rows), fmt[0], f"{fmt[1]} .. {fmt[2]}"))
        width = [max(len(r[i]) for r in [head, *body]) for i in range(len(head))]

        def row(cells):
            return "  ".join(c.ljust(wd) for c, wd in zip(cells, width)).rstrip()

        out = [f"{self.interpolator_type}: {self.total_rows} rows, "
               f"{len(body)} active families", row(head)]
        out += [row(r) for r in body] or ["(no active family)"]
        rc = self.region_coverage
        if rc is not None:
            out.append(f"active region: {rc.active_region_nodes} of {rc.total_support_nodes} "
                       f"nodes ({100.0 * rc.active_fraction:.1f}%)")
        if self.outside_model_points:
            out.append("points outside the model: " + ", ".join(
                f"{k}={v}" for k, v in sorted(self.outside_model_points.items())))
        return "\n".join(out)
