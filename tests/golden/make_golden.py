"""Writes tests/golden/known_answers.json: hand-derived known answers of the MEX loops.

No reference code can run here (the MEX files need MATLAB's mex.h/libmx, which the
image lacks), so these vectors come from the operator definitions worked by hand:

case das_single_pixel (forwardDelayAndSumMex.cpp:73-101):
    one pixel, tau = 10.25, R = 2.5, nT = 20, image value 2.0.
    taps r = ceil(7.75)..floor(12.75) = 8..12, weights 1 - |r - 10.25|/2.5
    = 0.1, 0.5, 0.9, 0.7, 0.3  ->  y[8..12] = 0.2, 1.0, 1.8, 1.4, 0.6.
case das_edge_drop: tau = 0.5, R = 2.0, nT = 4: taps r = -1..2; r = -1 is dropped;
    weights r=0: 0.75, r=1: 0.75, r=2: 0.25.
case sir_single_pixel (forwardMex.cpp:118-159):
    Ls = 5 (floor(Ls/2) = 2), one SIR column [0, 1, 2, 3, 0], window [2, 4] (1-based),
    tau = 10, R = 1.5: hat taps r = 9, 10, 11 with weights 1/3, 1, 1/3;
    t = r + s - 2 for s = 1, 2, 3 (0-based): y[8] = 1/3, y[9] = 1 + 2/3,
    y[10] = 1/3 + 2 + 1, y[11] = 2/3 + 3, y[12] = 1.
case das_transpose_single_detector (transposeDelayAndSumMex.cpp:67-96):
    y = [0, 1, 2, ..., 9] (nT = 10), tau = 4.4, R = 2: taps r = 3..6, weights
    0.3, 0.8, 0.7, 0.2 -> x = 0.3*3 + 0.8*4 + 0.7*5 + 0.2*6 = 8.8.
"""
import json
import os

cases = {
    "das_single_pixel": {"tau": 10.25, "R": 2.5, "nT": 20, "value": 2.0,
                         "expect": {"8": 0.2, "9": 1.0, "10": 1.8, "11": 1.4, "12": 0.6}},
    "das_edge_drop": {"tau": 0.5, "R": 2.0, "nT": 4, "value": 1.0,
                      "expect": {"0": 0.75, "1": 0.75, "2": 0.25, "3": 0.0}},
    "sir_single_pixel": {"tau": 10.0, "R": 1.5, "nT": 16, "value": 1.0, "Ls": 5,
                         "sir": [0.0, 1.0, 2.0, 3.0, 0.0], "window": [2, 4],
                         "expect": {"8": 1.0 / 3.0, "9": 5.0 / 3.0, "10": 10.0 / 3.0,
                                    "11": 11.0 / 3.0, "12": 1.0}},
    "das_transpose_single_detector": {"tau": 4.4, "R": 2.0, "nT": 10,
                                      "signal": list(range(10)), "expect": 8.8},
}

if __name__ == "__main__":
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json"), "w") as f:
        json.dump(cases, f, indent=1)
