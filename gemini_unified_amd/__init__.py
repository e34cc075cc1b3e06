"""MI355X-native Green-FK engine for GEMINI_UNIFIED's (omega, l) minors path."""
from .earthmodel import (NodeEarthmodel, ExternalRadialNodes, FrequencyGrid, global_model, global_model_json,
                         regional_model, regional_model_json, dll1_of, MC_PI, ATTMODES)
from .greenfk import GreenFKEngine, STAT_NAMES

__all__ = ["NodeEarthmodel", "ExternalRadialNodes", "FrequencyGrid", "global_model", "global_model_json", "dll1_of",
           "MC_PI", "ATTMODES", "GreenFKEngine", "STAT_NAMES", "regional_model", "regional_model_json"]
