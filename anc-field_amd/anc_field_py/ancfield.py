"""AncField: the reference wrapper's call surface over the afx C ABI.

Same constructor and methods as anc_field_py/ancfield.py:32-137 of the
reference (AncField(device, model, minsize), Pml, AddMic, AddSource,
Visualize, Run), but in-process through ctypes: no executable, no text
protocol and no temporary files.  Run(length) returns (inputs, outputs) as
numpy arrays, like the reference: inputs[nsrc, length*fs] and the receiver
traces (a 1-D array when there is a single receiver, as np.loadtxt gives).
"""
import os

import numpy as np

from . import _afx


class AncField:
    fs = 32000  # ancfield.py:34, main.cpp:118

    def __init__(self, device, model=None, minsize=None, exeloc=None):
        if device != "gpu" and device != "cpu":
            raise Exception("Device is either 'gpu' or 'cpu'")
        if device == "cpu":
            raise Exception("this build runs on HIP devices only")
        if model is not None and not os.path.exists(os.path.join(model, "model.xml")):
            raise Exception(model + " is not a valid model location")
        self.model = model
        self.minsize = minsize
        self.mics = []
        self.inputs = []
        self.pos_inputs = []
        self.pml_width = 0
        self.vis_level = None

    def Pml(self, width):
        self.pml_width = int(width)

    def AddMic(self, position):
        self.mics += [position]

    def AddSource(self, pos, data):
        self.inputs.append(np.asarray(data))
        self.pos_inputs.append(pos)

    def Visualize(self, level):
        # the live MJPEG view of the reference is not part of this build; the
        # request is recorded and ignored
        self.vis_level = level

    def Run(self, length):
        steps = int(np.float32(length) * np.float32(self.fs))  # (int)(sec * 32000.f), main.cpp:119-121
        for e in self.inputs:
            # the reference's check, kept as written (ancfield.py:81-83): shorter
            # sources are zero-padded by the simulation
            if np.size(e) < length / self.fs:
                raise Exception("Input shorter than simulation interval")
        sim = _afx.Sim()  # reference constants: c=340, dx=0.02, dt=3.125e-5, compat mode
        try:
            if self.model is not None:
                sim.load_kerkythea(os.path.join(self.model, "model.xml"))
            if self.minsize is not None:
                sim.set_min_size(*[float(v) for v in self.minsize])
            if self.pml_width:
                sim.set_pml(self.pml_width)
            for pos, data in zip(self.pos_inputs, self.inputs):
                sim.add_source([float(v) for v in pos], np.asarray(data, np.float32))
            for m in self.mics:
                sim.add_receiver([float(v) for v in m])
            sim.run(steps)
            out = sim.receivers()
        finally:
            sim.close()
        n = int(length * self.fs)
        input_data = np.zeros(n)
        for i in self.inputs:
            input_data = np.vstack([input_data, np.asarray(i, np.float64)[0:n]])
        outputs = out[0] if out.shape[0] == 1 else out
        return np.atleast_2d(input_data)[1:, :], outputs
