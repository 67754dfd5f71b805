"""Dispersive material fits."""

from .debye import DebyeFit, djordjevic_sarkar_eps, fit_callable, fit_debye, fit_djordjevic_sarkar

__all__ = ["DebyeFit", "djordjevic_sarkar_eps", "fit_callable", "fit_debye", "fit_djordjevic_sarkar"]
