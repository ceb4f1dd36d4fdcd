"""CPU checks of the denoiser's network description (aph_denoiser_arch, no GPU):
the parameter layout tiles the vector exactly, every op's shapes agree (a
convolution reads the whole of its input buffer and writes within its output,
pooling halves and upsampling doubles the level), every buffer slice is written
before it is read, and the backward schedule's accumulation rule holds: each
concatenation buffer is fully written by one convolution's input gradient
before the pooling gradient adds into its low channels."""
import numpy as np

from aurora_amd import capi
from aurora_amd.denoiser import arch, param_count


def test_parameter_layout_tiles_the_vector():
    layers, bufs, ops = arch()
    spans = []
    for L in layers:
        assert L.cin <= L.cin_padded and L.cout <= L.cout_padded
        assert L.cin_padded % 32 == 0 and L.cout_padded % 16 == 0  # MFMA K step and N tile
        spans.append((L.weight_offset, 9 * L.cout_padded * L.cin_padded))
        spans.append((L.bias_offset, L.cout_padded))
    spans.sort()
    pos = 0
    for off, n in spans:
        assert off == pos
        pos += n
    assert pos == param_count()
    assert 1_000_000 < pos < 2_000_000


def test_schedule_shapes_and_order():
    layers, bufs, ops = arch()
    nb = len(bufs)
    out_buf = nb - 1
    written = {b: np.zeros(bufs[b].channels, dtype=bool) for b in range(nb)}
    written[0][:] = True  # packed input
    convs = 0
    for op in ops:
        if op.kind == capi.APH_DN_OP_CONV:
            L = layers[op.layer]
            convs += 1
            assert bufs[L.in_buffer].channels == L.cin_padded
            assert written[L.in_buffer].all(), L.name
            assert bufs[L.in_buffer].level == bufs[L.out_buffer].level
            assert L.out_offset + L.cout_padded <= bufs[L.out_buffer].channels
            assert not written[L.out_buffer][L.out_offset:L.out_offset + L.cout].any()
            written[L.out_buffer][L.out_offset:L.out_offset + L.cout] = True
            assert L.relu == (L.out_buffer != out_buf)
        else:
            src, dst = bufs[op.src], bufs[op.dst]
            assert written[op.src][op.src_offset:op.src_offset + op.channels].all()
            assert op.dst_offset + op.channels <= dst.channels
            step = 1 if op.kind == capi.APH_DN_OP_POOL else -1
            assert dst.level == src.level + step
            written[op.dst][op.dst_offset:op.dst_offset + op.channels] = True
    assert convs == len(layers)
    assert all(w.all() for b, w in written.items() if b != out_buf)


def test_backward_accumulation_rule():
    # walking the schedule backwards: a convolution's input gradient overwrites
    # its whole input buffer; pooling gradients add into a slice that must
    # already be written; upsampling gradients overwrite their source buffer.
    layers, bufs, ops = arch()
    written = {b: np.zeros(bufs[b].channels, dtype=bool) for b in range(len(bufs))}
    for op in reversed(ops):
        if op.kind == capi.APH_DN_OP_CONV:
            L = layers[op.layer]
            if op.layer > 0:
                assert not written[L.in_buffer].any()
                written[L.in_buffer][:] = True
        elif op.kind == capi.APH_DN_OP_POOL:
            assert written[op.src][op.src_offset:op.src_offset + op.channels].all()
        else:
            assert not written[op.src].any()
            written[op.src][:] = True
