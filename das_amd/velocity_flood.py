"""Flood-model container consumed by the RTM engine.

``FloodModels`` has the fields of tsw_rtm/velocity_flood.py:15-21; the RTM
path reads only ``sediment_flood`` (forward model) and ``rock_flood``
(back-propagation model), rtm_engine.py:78 and :95.  Building flood models is
host-side model preparation outside the migration path.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class FloodModels:
    true_model: np.ndarray
    sediment_flood: np.ndarray
    rock_flood: np.ndarray
    flood_line_sed: int
    flood_line_rock: int


def flood_models_from(migration_model: np.ndarray, rock_model: np.ndarray = None) -> FloodModels:
    """FloodModels with one migration model used for both propagations
    (the single-model RTM case), or distinct forward / backward models."""
    m = np.ascontiguousarray(migration_model, dtype=np.float32)
    r = m if rock_model is None else np.ascontiguousarray(rock_model, dtype=np.float32)
    return FloodModels(true_model=m, sediment_flood=m, rock_flood=r, flood_line_sed=0,
                       flood_line_rock=0)
