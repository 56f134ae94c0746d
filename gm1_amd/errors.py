"""Refusal convention at the operator boundary.

Mirrors the reference's ``RefusalDetail``/``CM1Refusal``/``refuse``
(``src/cm1core/errors.py:72-82,235-270``): an unsupported physics branch stops
before any launch with a named code, gate, option, value, reason and remedy.
ABI misuse (wrong dtype, shape, device, aliasing) raises ``TypeError`` or
``ValueError`` instead, as the reference's launchers do.
"""

from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class RefusalDetail:
    code: str
    gate: str
    reason: str
    remedy: str
    option: str | None = None
    value: object | None = None

    def as_dict(self) -> dict[str, object]:
        payload: dict[str, object] = {
            "status": "refused", "code": self.code, "gate": self.gate,
            "reason": self.reason, "remedy": self.remedy,
        }
        if self.option is not None:
            payload["option"] = self.option
            payload["value"] = self.value
        return payload


class CM1Refusal(RuntimeError):
    """An intentional stop that names the invalid execution path."""

    def __init__(self, detail: RefusalDetail):
        self.detail = detail
        message = f"{detail.code} [{detail.gate}]: {detail.reason}"
        if detail.option is not None:
            message += f" (option {detail.option}={detail.value!r})"
        message += f" Remedy: {detail.remedy}"
        super().__init__(message)


def refuse(*, code: str, gate: str, reason: str, remedy: str,
           option: str | None = None, value: object | None = None):
    raise CM1Refusal(RefusalDetail(code=code, gate=gate, reason=reason,
                                   remedy=remedy, option=option, value=value))
