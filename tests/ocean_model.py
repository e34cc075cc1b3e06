"""Test model with an ocean: the seeded 200-knot global model with its mantle compressed to end
3 km below the surface and a fluid ocean layer (3 knots, 1020 kg/m3, 1.5 km/s) on top."""
import numpy as np

from gemini_unified_amd import NodeEarthmodel
from gemini_unified_amd.earthmodel import global_model_json


def ocean_model(depth=3000.0, seed=1):
    js = global_model_json(seed, 200)
    R = js["earth_radius"]
    npp = js["node_properties"]
    L = js["num_layers"]
    idx = [i for i, lay in enumerate(npp["layer_index"]) if lay == L]
    r = R - np.asarray(npp["depth"], float)
    rm = r[idx]
    rm = rm[0] + (rm - rm[0]) * ((R - depth) - rm[0]) / (rm[-1] - rm[0])
    for k, i in enumerate(idx):
        npp["depth"][i] = float(R - rm[k])
    for d in (depth, depth / 2, 0.0):
        npp["layer_index"].append(L + 1)
        npp["depth"].append(d)
        npp["density"].append(1020.0)
        npp["P-velocity"].append(1500.0)
        npp["S-velocity"].append(0.0)
        npp["Qkappa"].append(57823.0)
        npp["Qmu"].append(-1.0)
    js["num_layers"] = L + 1
    js["num_nodes"] += 3
    js["radius_upper_layer_boundary_m"] = js["radius_upper_layer_boundary_m"][:-1] + [R - depth, R]
    js["isfluid"] = js["isfluid"] + [1]
    js["index_upper_boundary"] = js["index_upper_boundary"] + [js["num_nodes"]]
    return NodeEarthmodel.from_json(js)


def ocean_nodes(nem, d=1.0):
    """Outer-core and mantle nodes, an ocean-bottom receiver 500 m below the ocean, and nodes d
    metres either side of the ocean bottom and near the surface (1-based receiver index 3)."""
    from gemini_unified_amd import ExternalRadialNodes
    ruc = nem.rb[-2]
    rn = np.array([2000e3, 4000e3, ruc - 500.0, ruc - d, ruc + d, nem.rearth - 500.0])
    st = np.array([1, 0, 0, 0, 1, 1], np.int32)
    la = np.array([3, 4, 4, 4, 5, 5], np.int32)
    return ExternalRadialNodes(rearth=nem.rearth, rnod=rn, state=st, layer=la), 3
