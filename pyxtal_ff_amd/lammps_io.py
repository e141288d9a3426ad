"""The LAMMPS ML-IAP export files of a trained network: ``NN_weights.txt``
and ``DescriptorParam.txt``.

Same text format as the reference's ``NeuralNetwork.save_weights_to_txt``
(``pyxtal_ff/models/neuralnetwork.py:719-835``), so a potential trained with
this package loads in LAMMPS (``pair_style mliap model nn``) and a potential
exported by the reference loads here:

  * a header with the species count and the parameter count of one network;
  * ``NET ndescriptors nlayers act_1 width_1 ... act_L width_L`` (activation
    names lower case, the output layer included, written once);
  * per element: ``# Scale0`` (descriptor minima), ``# Scale1`` (max - min),
    five numbers per line at 13 decimals, then ``# Coefficients (El)``: for
    each layer, for each output neuron, the bias followed by its weights.
"""
import numpy as np
import torch

from . import potential as pot

_ACT_OUT = {v.lower(): v for v in pot.ACTIVATIONS}


def _fmt_block(values, fmt):
    """Numbers five to a line (a last partial line ends with a newline)."""
    out, n = [], len(values)
    for k, v in enumerate(values, 1):
        out.append(fmt.format(v) + " ")
        if k % 5 == 0 or k == n:
            out.append("\n")
    return "".join(out)


def read_nn_weights(path):
    """Parse NN_weights.txt -> dict(ndesc, layers [(activation, width)],
    elements {name: dict(scale0, scale1, coef)}, nparams)."""
    with open(path) as f:
        lines = [ln.strip() for ln in f]
    header, net, sections = None, None, {}
    block, target = None, None
    scales = {}
    for ln in lines:
        if not ln:
            continue
        if ln.startswith("#"):
            tag = ln[1:].strip()
            if tag.startswith("Scale0") or tag.startswith("Scale1"):
                block = tag[:6].lower()
                scales[block] = []
                target = scales[block]
            elif tag.startswith("Coefficients"):
                el = tag[tag.index("(") + 1: tag.index(")")]
                sections[el] = {"scale0": scales.pop("scale0", []), "scale1": scales.pop("scale1", []), "coef": []}
                target = sections[el]["coef"]
            continue
        tok = ln.split()
        if header is None:
            header = (int(tok[0]), int(tok[1]))
        elif tok[0] == "NET":
            nl = int(tok[2])
            net = (int(tok[1]), [(tok[3 + 2 * q], int(tok[4 + 2 * q])) for q in range(nl)])
        elif target is not None:
            target.extend(float(v) for v in tok)
    if header is None or net is None:
        raise ValueError(f"{path}: not an NN_weights file")
    nspec, nparams = header
    nd, layers = net
    for el, d in sections.items():
        for key in ("scale0", "scale1", "coef"):
            d[key] = np.asarray(d[key], dtype=np.float64)
        if len(d["scale0"]) != nd or len(d["scale1"]) != nd:
            raise ValueError(f"{path}: {el}: {len(d['scale0'])} scales for {nd} descriptors")
        if len(d["coef"]) != nparams:
            raise ValueError(f"{path}: {el}: {len(d['coef'])} coefficients, header says {nparams}")
    if len(sections) != nspec:
        raise ValueError(f"{path}: {len(sections)} element sections, header says {nspec}")
    return {"ndesc": nd, "layers": layers, "elements": sections, "nparams": nparams}


def potential_from_weights(w, device="cpu"):
    """NNPotential with the file's layers, coefficients and descriptor
    ranges (drange = [Scale0, Scale0 + Scale1])."""
    layers = w["layers"]
    hidden = [wd for _, wd in layers[:-1]]
    acts = [_ACT_OUT[a.lower()] for a, _ in layers]
    if layers[-1][1] != 1:
        raise ValueError("the last layer must have one output")
    els = list(w["elements"])
    drange = {e: np.stack([d["scale0"], d["scale0"] + d["scale1"]], 1) for e, d in w["elements"].items()}
    model = pot.NNPotential(els, w["ndesc"], hidden, activation=acts[:-1], out_activation=acts[-1],
                            drange=drange)
    # the file's Scale1 is max - min: keep 1/Scale1 exactly rather than
    # 1/((min + range) - min)
    for e, d in w["elements"].items():
        getattr(model, "scale_" + e).copy_(torch.as_tensor(1.0 / d["scale1"]))
        getattr(model, "range_" + e).copy_(torch.as_tensor(d["scale1"]))
    with torch.no_grad():
        for e, d in w["elements"].items():
            c = d["coef"]
            off = 0
            nin = w["ndesc"]
            for lin, (_, wd) in zip(model.models[e].layers, layers):
                blk = c[off: off + wd * (nin + 1)].reshape(wd, nin + 1)
                lin.bias.copy_(torch.as_tensor(blk[:, 0]))
                lin.weight.copy_(torch.as_tensor(blk[:, 1:]))
                off += wd * (nin + 1)
                nin = wd
    return model.to(device)


def write_nn_weights(path, model):
    """Write an NNPotential in the NN_weights.txt format."""
    els = model.elements
    m0 = model.models[els[0]]
    widths = [lin.out_features for lin in m0.layers]
    nparams = sum(p.numel() for p in m0.parameters())
    out = ["# Neural networks weights generated in PyXtal_FF \n", "# total_species nparams \n",
           f"{len(els)} {nparams} \n\n"]
    for k, e in enumerate(els):
        m = model.models[e]
        if k == 0:
            out.append("# NET ndescriptors nlayers activation_func for layer 1, number of nodes for layer 1, "
                       "..., ..., ...., \n")
            out.append(f"NET {model.nin} {len(widths)} ")
            for act, wd in zip(m.acts, widths):
                out.append(f"{act.lower()} {wd} ")
            out.append("\n\n")
        dmin = getattr(model, "dmin_" + e).detach().cpu().numpy()
        drng = getattr(model, "range_" + e).detach().cpu().numpy()
        out.append("# Scale0 \n" + _fmt_block(dmin, "{:20.13f}") + "\n")
        out.append("# Scale1 \n" + _fmt_block(drng, "{:20.13f}") + "\n")
        out.append(f"# Coefficients ({e})\n")
        vals = []
        for lin in m.layers:
            blk = torch.cat([lin.bias.detach()[:, None], lin.weight.detach()], 1).cpu().numpy()
            vals.extend(blk.reshape(-1))
        cnt = 1
        for v in vals:
            out.append("{:18.13f} \n".format(v) if cnt % 5 == 0 else "{:18.13f} ".format(v))
            cnt += 1
        out.append("\n" if (cnt - 1) % 5 == 0 else "\n\n")
    with open(path, "w") as f:
        f.write("".join(out))


def read_descriptor_param(path):
    """DescriptorParam.txt -> dict of its keys (numbers converted, element
    lists as lists)."""
    d = {}
    with open(path) as f:
        for ln in f:
            ln = ln.strip()
            if not ln or ln.startswith("#"):
                continue
            tok = ln.split()
            key, vals = tok[0], tok[1:]
            conv = []
            for v in vals:
                try:
                    conv.append(int(v))
                except ValueError:
                    try:
                        conv.append(float(v))
                    except ValueError:
                        conv.append(v)
            d[key] = conv if key in ("elems", "radelems", "welems") else (conv[0] if len(conv) == 1 else conv)
    return d


def write_descriptor_param(path, des_info, elements, atomic_numbers=None):
    """DescriptorParam.txt from the reference's des_info dict
    (type, Rc, parameters, weights) -- neuralnetwork.py:795-831."""
    t = des_info["type"]
    out = ["# Descriptor parameters generated in PyXtal_FF \n\n", "# Required \n", f"rcutfac {des_info['Rc']} \n"]
    if t in ("SO3", "SOAP"):
        p = des_info["parameters"]
        out += [f"nmax {p['nmax']} \n", f"lmax {p['lmax']} \n", f"alpha {p['alpha']} \n\n"]
    else:
        out.append(f"twojmax {des_info['parameters']['lmax'] * 2} \n\n")
    out += ["# Elements \n\n", f"nelems {len(elements)} \n", "elems " + "".join(f"{e} " for e in elements) + "\n"]
    if t in ("snap", "SNAP", "SO3", "SOAP"):
        out.append("radelems " + "".join("0.5 " for _ in elements) + "\n")
        if t in ("snap", "SNAP"):
            out.append("welems " + "".join(f"{des_info['weights'][e]} " for e in elements) + "\n\n")
        else:
            if atomic_numbers is None:
                raise ValueError("SO3 export needs the atomic numbers of the elements")
            out.append("welems " + "".join(f"{z} " for z in atomic_numbers) + "\n\n")
    if t in ("snap", "SNAP"):
        out += [f"rfac0 {des_info['parameters']['rfac']} \n", "rmin0 0 ", "\n", "switchflag 1 \n", "bzeroflag 0 \n"]
    with open(path, "w") as f:
        f.write("".join(out))
