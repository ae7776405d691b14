"""Setup diagnostics report with the reference's fields (LI/_diagnostics.py:7-102)."""
from __future__ import annotations

from dataclasses import dataclass, field


@dataclass(frozen=True)
class ConstraintFamilyDiagnostics:
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
class RegionCoverageDiagnostics:
    total_support_nodes: int
    active_region_nodes: int
    inactive_region_nodes: int
    active_fraction: float


@dataclass(frozen=True)
class ConstraintDiagnosticsReport:
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
        rc = self.region_coverage
        return {
            "interpolator_type": self.interpolator_type,
            "total_rows": self.total_rows,
            "families": {n: {"active": f.active, "row_count": f.row_count,
                             "dropped_rows": f.dropped_rows,
                             "effective_weight_mean": f.effective_weight_mean,
                             "effective_weight_min": f.effective_weight_min,
                             "effective_weight_max": f.effective_weight_max,
                             "source_point_count": f.source_point_count,
                             "outside_model_point_count": f.outside_model_point_count}
                         for n, f in self.families.items()},
            "region_coverage": None if rc is None else {
                "total_support_nodes": rc.total_support_nodes,
                "active_region_nodes": rc.active_region_nodes,
                "inactive_region_nodes": rc.inactive_region_nodes,
                "active_fraction": rc.active_fraction},
            "outside_model_points": dict(self.outside_model_points),
        }

    def summary(self) -> str:
        lines = [f"Constraint diagnostics for {self.interpolator_type}",
                 f"Total rows: {self.total_rows}", "Active families:"]
        active = self.active_families
        if not active:
            lines.append("  - none")
        for name, f in sorted(active.items()):
            dropped = "unknown" if f.dropped_rows is None else str(f.dropped_rows)
            mw = "n/a" if f.effective_weight_mean is None else f"{f.effective_weight_mean:.4g}"
            lines.append(f"  - {name}: rows={f.row_count}, dropped={dropped}, mean_weight={mw}")
        if self.region_coverage is not None:
            rc = self.region_coverage
            lines.append(f"Region coverage: {rc.active_region_nodes}/{rc.total_support_nodes} "
                         f"({rc.active_fraction:.1%})")
        if self.outside_model_points:
            lines.append("Outside-model points:")
            for k, v in sorted(self.outside_model_points.items()):
                lines.append(f"  - {k}: {v}")
        return "\n".join(lines)
