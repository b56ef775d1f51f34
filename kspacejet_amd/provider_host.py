"""A minimal KSpaceJet Provider host in Python (ctypes), for driving the MI355X
provider modules without the reference runtime.

It plays the host's side of the Provider ABI (include/ksj_provider_abi.h, the
layout of libs/recon/kspacejet-provider-sdk/include/kspacejet/provider/provider.h):
it loads a module, calls ``ksj_provider_query``, creates operators, contexts and
key states through the lifecycle table, and runs firings through a firing lease
whose callbacks it implements.  The lease records every call the operator makes
(the counters the reference's provider unit tests assert on,
tests/unit/providers/kspacejet-noncartesian-recon/noncartesian_recon_provider_tests.cpp:74-412)
and holds its payloads in 64-byte-aligned host buffers.

    mod = ProviderModule("noncartesian_recon")
    st, op = mod.create("radial_gridding_reconstruct", config_json)
    lease = Lease(inputs=[...], output=OutputSpec(...), scratch_bytes=...)
    st, result, message = op.process(lease)
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import json
import os
from typing import List, Optional, Sequence

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
TYPES_JSON = os.path.join(PKG, "csrc", "provider", "types.json")

u32, u64, vp = C.c_uint32, C.c_uint64, C.c_void_p

# ---- status and enum values (provider.h:98-209) ----
OK, BAD_ABI, INVALID_ARGUMENT, UNSUPPORTED, RESOURCE_EXHAUSTED = 0, 1, 2, 3, 4
CANCELLED, FAILED_PRECONDITION, INTERNAL_ERROR, CONTRACT_VIOLATION = 5, 6, 7, 8
PROCESS_DONE = 0
SCAN_END_NORMAL, SCAN_END_CANCELLED, SCAN_END_FAILED = 0, 1, 2
MEMORY_HOST_PAGEABLE, MEMORY_HOST_PINNED, MEMORY_DEVICE = 1, 2, 4
LEASE_CAP_INPUT_BATCHES, LEASE_CAP_OUTPUT_GRANTS, LEASE_CAP_SCRATCH = 1, 2, 4
SERIAL_INSTANCE = 0

_KIND = {"buffer_handle": 0, "message_handle": 1, "control_token": 2, "opaque_handle": 3}
_ELEMENT = {"float32": 6, "float64": 7, "complex_float32": 9, "complex_float64": 10}
_LAYOUT = {"canonical_contiguous": 1 << 0, "channel_major_contiguous": 1 << 1, "row_major_contiguous": 1 << 2,
           "column_major_contiguous": 1 << 3, "opaque": 1 << 4}
_STRIDES = {"canonical": 1 << 16, "explicit_byte": 1 << 17}
_DOMAIN = {"host_normal": MEMORY_HOST_PAGEABLE}
_MUTABILITY = {"immutable_after_publish": 0, "mutable_exclusive": 1}


# ---- ABI structs ----

class AbiHeader(C.Structure):
    _fields_ = [("struct_size", u32), ("reserved0", u32), ("capability_bits", u64), ("reserved", u64 * 2)]


def header(cls, caps: int = 0) -> AbiHeader:
    return AbiHeader(C.sizeof(cls), 0, caps, (u64 * 2)(0, 0))


class Utf8View(C.Structure):
    _fields_ = [("abi", AbiHeader), ("data", vp), ("size", u64)]


class ByteView(C.Structure):
    _fields_ = [("abi", AbiHeader), ("data", vp), ("size", u64)]


class Digest256(C.Structure):
    _fields_ = [("abi", AbiHeader), ("bytes", C.c_uint8 * 32)]


class ErrorView(C.Structure):
    _fields_ = [("abi", AbiHeader), ("status", u32), ("category", u32), ("provider_error_code", u64),
                ("message", Utf8View)]


class TypeDescriptorView(C.Structure):
    _fields_ = [("abi", AbiHeader), ("type_ref", Utf8View), ("type_identity_digest", Digest256),
                ("payload_kind", u32), ("element_type", u32), ("rank", u32), ("dimension_names", C.POINTER(Utf8View)),
                ("layout_flags", u64), ("stride_bytes", u64 * 8), ("allowed_memory_domains", u32),
                ("minimum_alignment", u32), ("mutability", u32), ("reserved0", u32)]


class OperatorDescriptor(C.Structure):
    _fields_ = [("abi", AbiHeader), ("operator_id", Utf8View), ("max_in_flight", u32), ("thread_safety", u32),
                ("max_private_threads", u32), ("max_input_items_per_firing", u32),
                ("max_output_items_per_firing", u32), ("max_output_bytes_per_firing", u64),
                ("max_scratch_bytes_per_firing", u64), ("max_retained_input_bytes", u64),
                ("max_async_tail_bytes", u64)]


class ProviderDescriptor(C.Structure):
    _fields_ = [("abi", AbiHeader), ("provider_id", Utf8View), ("bundle_digest", Digest256),
                ("operator_count", u32), ("reserved0", u32), ("operators", C.POINTER(OperatorDescriptor))]


class QueryRequest(C.Structure):
    _fields_ = [("abi", AbiHeader), ("host_capability_bits", u64), ("host_build_id", Utf8View)]


class HostAllocationRequest(C.Structure):
    _fields_ = [("abi", AbiHeader), ("byte_count", u64), ("alignment", u32), ("memory_domain", u32),
                ("resource_occurrence_id", u64), ("allocation_flags", u64)]


class MutablePayloadView(C.Structure):
    _fields_ = [("abi", AbiHeader), ("data", vp), ("capacity_bytes", u64), ("committed_bytes", u64),
                ("memory_domain", u32), ("alignment", u32), ("type", TypeDescriptorView)]


class HostServices(C.Structure):
    _fields_ = [("abi", AbiHeader), ("host_context", vp), ("allocate", vp), ("map_allocation", vp),
                ("release_allocation", vp), ("backend_executor", vp)]


class OperatorCreateRequest(C.Structure):
    _fields_ = [("abi", AbiHeader), ("operator_id", Utf8View), ("canonical_config", ByteView),
                ("host_services", vp)]


class ExecutionContextDescriptor(C.Structure):
    _fields_ = [("abi", AbiHeader), ("numa_node", u32), ("device_ordinal", u32), ("execution_context_id", u64),
                ("resource_domain_id", u64), ("max_backend_concurrency", u64), ("host_services", vp)]


class KeyStateDescriptor(C.Structure):
    _fields_ = [("abi", AbiHeader), ("semantic_key", ByteView), ("placement_key", u64),
                ("key_state_generation", u64), ("home_shard", u64)]


class ScanStartDescriptor(C.Structure):
    _fields_ = [("abi", AbiHeader), ("run_id", Utf8View), ("scan_id", Utf8View),
                ("normalized_scan_facts_digest", Digest256), ("execution_plan_digest", Digest256),
                ("terminal_epoch", u64)]


class PayloadView(C.Structure):
    _fields_ = [("abi", AbiHeader), ("data", vp), ("byte_count", u64), ("memory_domain", u32), ("alignment", u32),
                ("type", TypeDescriptorView)]


class InputItemView(C.Structure):
    _fields_ = [("abi", AbiHeader), ("payload", PayloadView), ("metadata", ByteView), ("semantic_key_hash", u64),
                ("order_key", u64), ("item_ordinal", u64)]


class InputBatchView(C.Structure):
    _fields_ = [("abi", AbiHeader), ("items", C.POINTER(InputItemView)), ("item_count", u32), ("input_port", u32),
                ("batch_id", u64), ("order_domain", u64)]


class ScratchView(C.Structure):
    _fields_ = [("abi", AbiHeader), ("data", vp), ("byte_count", u64), ("memory_domain", u32), ("alignment", u32),
                ("resource_occurrence_id", u64)]


class KeyStateView(C.Structure):
    _fields_ = [("abi", AbiHeader), ("data", vp), ("byte_count", u64), ("key_state_generation", u64),
                ("slot_generation", u64)]


class FiringLeaseInfo(C.Structure):
    _fields_ = [("abi", AbiHeader), ("resource_occurrence_id", u64), ("slot_generation", u64),
                ("terminal_epoch", u64), ("input_batch_count", u32), ("output_grant_count", u32),
                ("reserved_output_bytes", u64), ("reserved_scratch_bytes", u64)]


class RetentionRequest(C.Structure):
    _fields_ = [("abi", AbiHeader), ("requested_bytes", u64), ("retention_class", u64),
                ("maximum_lifetime_epoch", u64)]


class AsyncRegistration(C.Structure):
    _fields_ = [("abi", AbiHeader), ("resource_occurrence_id", u64), ("slot_generation", u64),
                ("terminal_epoch", u64), ("reserved_output_bytes", u64), ("reserved_scratch_bytes", u64),
                ("completion_deadline_ns", u64), ("maximum_completion_callbacks", u32), ("reserved0", u32)]


class AsyncCompletion(C.Structure):
    _fields_ = [("abi", AbiHeader), ("outcome", u32), ("sealed_output_count", u32), ("completion_sequence", u64),
                ("terminal_epoch", u64)]


class CancellationView(C.Structure):
    _fields_ = [("abi", AbiHeader), ("state", u32), ("reserved0", u32), ("terminal_epoch", u64),
                ("cancellation_generation", u64)]


class OutputSealDescriptor(C.Structure):
    _fields_ = [("abi", AbiHeader), ("output_port", u32), ("produced_item_count", u32),
                ("produced_byte_count", u64), ("semantic_key_hash", u64), ("order_key", u64),
                ("type", TypeDescriptorView), ("metadata", ByteView)]


E = C.POINTER(ErrorView)
MapFn = C.CFUNCTYPE(u32, vp, vp, C.POINTER(MutablePayloadView), E)
SealFn = C.CFUNCTYPE(u32, vp, vp, C.POINTER(OutputSealDescriptor), E)
ReleaseFn = C.CFUNCTYPE(u32, vp, vp, E)
GetInfoFn = C.CFUNCTYPE(u32, vp, vp, C.POINTER(FiringLeaseInfo), E)
GetInputBatchFn = C.CFUNCTYPE(u32, vp, vp, u32, C.POINTER(InputBatchView), E)
GetScratchFn = C.CFUNCTYPE(u32, vp, vp, C.POINTER(ScratchView), E)
AcquireOutputFn = C.CFUNCTYPE(u32, vp, vp, u32, C.POINTER(vp), E)


class OutputGrantCallbacks(C.Structure):
    _fields_ = [("abi", AbiHeader), ("host_context", vp), ("map_mutable_payload", MapFn), ("seal", SealFn),
                ("release", ReleaseFn)]


class FiringLeaseCallbacks(C.Structure):
    _fields_ = [("abi", AbiHeader), ("host_context", vp), ("output_grants", C.POINTER(OutputGrantCallbacks)),
                ("get_info", GetInfoFn), ("get_input_batch", GetInputBatchFn), ("get_scratch", GetScratchFn),
                ("get_key_state", vp), ("acquire_output_grant", AcquireOutputFn), ("retain_input", vp),
                ("release_retention", vp), ("register_async", vp), ("complete_async", vp), ("release_async", vp),
                ("get_cancellation", vp)]


class ProcessResult(C.Structure):
    _fields_ = [("abi", AbiHeader), ("outcome", u32), ("sealed_output_count", u32),
                ("consumed_input_item_count", u64), ("terminal_epoch", u64), ("async_token", vp)]


class ScanEndDescriptor(C.Structure):
    _fields_ = [("abi", AbiHeader), ("kind", u32), ("reserved0", u32), ("terminal_epoch", u64),
                ("completed_input_item_count", u64)]


class CancelContext(C.Structure):
    _fields_ = [("abi", AbiHeader), ("kind", u32), ("reserved0", u32), ("terminal_epoch", u64),
                ("cancellation_generation", u64), ("reason", Utf8View)]


OperatorCreateFn = C.CFUNCTYPE(u32, C.POINTER(OperatorCreateRequest), C.POINTER(vp), E)
ExecutionContextCreateFn = C.CFUNCTYPE(u32, vp, C.POINTER(ExecutionContextDescriptor), C.POINTER(vp), E)
KeyStateInitFn = C.CFUNCTYPE(u32, vp, vp, C.POINTER(KeyStateDescriptor), C.POINTER(vp), E)
OnStartFn = C.CFUNCTYPE(u32, vp, vp, vp, C.POINTER(ScanStartDescriptor), E)
ProcessBatchFn = C.CFUNCTYPE(u32, vp, vp, vp, vp, C.POINTER(FiringLeaseCallbacks), C.POINTER(ProcessResult), E)
OnScanEndFn = C.CFUNCTYPE(u32, vp, vp, vp, C.POINTER(ScanEndDescriptor), vp, C.POINTER(FiringLeaseCallbacks),
                          C.POINTER(ProcessResult), E)
OnCancelFn = C.CFUNCTYPE(u32, vp, vp, vp, C.POINTER(CancelContext), E)
KeyStateResetFn = C.CFUNCTYPE(None, vp, vp, vp)
ContextDestroyFn = C.CFUNCTYPE(None, vp, vp)
OperatorDestroyFn = C.CFUNCTYPE(None, vp)


class ProviderApi(C.Structure):
    _fields_ = [("abi", AbiHeader), ("operator_create", OperatorCreateFn),
                ("execution_context_create", ExecutionContextCreateFn), ("key_state_init", KeyStateInitFn),
                ("operator_on_start", OnStartFn), ("operator_process_batch", ProcessBatchFn),
                ("operator_on_scan_end", OnScanEndFn), ("operator_on_cancel", OnCancelFn),
                ("key_state_reset", KeyStateResetFn), ("execution_context_destroy", ContextDestroyFn),
                ("operator_destroy", OperatorDestroyFn)]


# Every ABI struct of the mirror by its C name (for layout checks).
STRUCTS = {
    "ksj_provider_abi_header": AbiHeader, "ksj_utf8_view": Utf8View, "ksj_byte_view": ByteView,
    "ksj_digest256": Digest256, "ksj_error_view": ErrorView, "ksj_type_descriptor_view": TypeDescriptorView,
    "ksj_operator_descriptor": OperatorDescriptor, "ksj_provider_descriptor": ProviderDescriptor,
    "ksj_provider_query_request": QueryRequest, "ksj_host_allocation_request": HostAllocationRequest,
    "ksj_mutable_payload_view": MutablePayloadView, "ksj_host_services": HostServices,
    "ksj_operator_create_request": OperatorCreateRequest,
    "ksj_execution_context_descriptor": ExecutionContextDescriptor,
    "ksj_key_state_descriptor": KeyStateDescriptor, "ksj_scan_start_descriptor": ScanStartDescriptor,
    "ksj_payload_view": PayloadView, "ksj_input_item_view": InputItemView, "ksj_input_batch_view": InputBatchView,
    "ksj_scratch_view": ScratchView, "ksj_key_state_view": KeyStateView, "ksj_firing_lease_info": FiringLeaseInfo,
    "ksj_retention_request": RetentionRequest, "ksj_async_registration": AsyncRegistration,
    "ksj_async_completion": AsyncCompletion, "ksj_cancellation_view": CancellationView,
    "ksj_output_seal_descriptor": OutputSealDescriptor, "ksj_output_grant_callbacks": OutputGrantCallbacks,
    "ksj_firing_lease_callbacks": FiringLeaseCallbacks, "ksj_process_result": ProcessResult,
    "ksj_scan_end_descriptor": ScanEndDescriptor, "ksj_cancel_context": CancelContext,
    "ksj_provider_api": ProviderApi,
}


# ---- helpers ----

_keep = []  # storage that borrowed views point at for the life of the process (type tables, texts)


def utf8(text: str) -> Utf8View:
    b = C.create_string_buffer(text.encode())
    _keep.append(b)
    return Utf8View(header(Utf8View), C.cast(b, vp), len(text.encode()))


def byte_view(data: Optional[bytes] = None) -> ByteView:
    if not data:
        return ByteView(header(ByteView), None, 0)
    b = C.create_string_buffer(data, len(data))
    _keep.append(b)
    return ByteView(header(ByteView), C.cast(b, vp), len(data))


def empty_error() -> ErrorView:
    e = ErrorView()
    e.abi = header(ErrorView)
    e.message.abi = header(Utf8View)
    return e


def error_text(e: ErrorView) -> str:
    if not e.message.data or not e.message.size:
        return ""
    return C.string_at(e.message.data, e.message.size).decode(errors="replace")


def view_text(v: Utf8View) -> str:
    return C.string_at(v.data, v.size).decode() if v.data else ""


def aligned_empty(nbytes: int, dtype=np.uint8, alignment: int = 64) -> np.ndarray:
    """A zeroed host array whose data pointer is `alignment`-aligned."""
    raw = np.zeros(nbytes + alignment, dtype=np.uint8)
    off = (-raw.ctypes.data) % alignment
    return raw[off:off + nbytes].view(dtype)


def aligned_copy(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    out = aligned_empty(a.nbytes, a.dtype).reshape(a.shape)
    out[...] = a
    return out


_types = None


def type_descriptor(type_ref: str) -> TypeDescriptorView:
    """The registry descriptor of `type_ref` as a host would place it in a payload view."""
    global _types
    if _types is None:
        _types = {t["type_ref"]: t for t in json.load(open(TYPES_JSON))}
    t = _types[type_ref]
    d = TypeDescriptorView()
    d.abi = header(TypeDescriptorView)
    d.type_ref = utf8(t["type_ref"])
    d.type_identity_digest.abi = header(Digest256)
    d.type_identity_digest.bytes[:] = list(bytes.fromhex(t["digest"]))
    d.payload_kind = _KIND[t["kind"]]
    d.element_type = _ELEMENT[t["element_type"]]
    d.rank = t["rank"]
    names = (Utf8View * max(1, t["rank"]))(*[utf8(n) for n in t["dimensions"]])
    _keep.append(names)
    d.dimension_names = C.cast(names, C.POINTER(Utf8View))
    d.layout_flags = _LAYOUT[t["layout"]] | _STRIDES[t["strides"]]
    d.allowed_memory_domains = 0
    for dom in t["allowed_domains"]:
        d.allowed_memory_domains |= _DOMAIN[dom]
    d.minimum_alignment = t["min_alignment_bytes"]
    d.mutability = _MUTABILITY[t["mutability"]]
    return d


# ---- firing lease ----

def _address(data) -> int:
    """Address of a payload: a numpy array (host) or a torch tensor (host or device)."""
    return data.data_ptr() if hasattr(data, "data_ptr") else data.ctypes.data


def _nbytes(data) -> int:
    return data.numel() * data.element_size() if hasattr(data, "data_ptr") else data.nbytes


@dataclasses.dataclass
class Item:
    """One input item: a payload (numpy array, or a device tensor with
    memory_domain=MEMORY_DEVICE) with its registry type and identity keys."""
    data: object
    type_ref: str
    semantic_key_hash: int = 71
    order_key: int = 73
    item_ordinal: int = 0
    metadata: bytes = b""
    byte_count: Optional[int] = None  # override of data.nbytes
    alignment: int = 64
    memory_domain: int = MEMORY_HOST_PAGEABLE
    type: Optional[TypeDescriptorView] = None  # override of the registry descriptor


@dataclasses.dataclass
class Batch:
    port: int
    items: List[Item]


@dataclasses.dataclass
class OutputSpec:
    nbytes: int
    type_ref: str
    dtype: type = np.complex64
    alignment: int = 64
    memory_domain: int = MEMORY_HOST_PAGEABLE
    type: Optional[TypeDescriptorView] = None


class Lease:
    """Host side of one firing: input batches, scratch, one output grant, and
    counters of every callback the operator makes."""

    def __init__(self, batches: Sequence[Batch], output: OutputSpec, scratch_bytes: int = 0,
                 terminal_epoch: int = 41, caps: int = LEASE_CAP_INPUT_BATCHES | LEASE_CAP_OUTPUT_GRANTS |
                 LEASE_CAP_SCRATCH):
        self.batches = list(batches)
        self.output_spec = output
        if output.memory_domain == MEMORY_DEVICE:  # a device-resident output grant
            import torch
            self.output = torch.empty(output.nbytes, dtype=torch.uint8, device="cuda")
        else:
            self.output = aligned_empty(output.nbytes, output.dtype)
        self.scratch = aligned_empty(max(scratch_bytes, 1))
        self.scratch_bytes = scratch_bytes
        self.terminal_epoch = terminal_epoch
        self.caps = caps
        self.input_batch_count: Optional[int] = None  # override of len(batches) in get_info
        self.reserved_output_bytes: Optional[int] = None
        self.reserved_scratch_bytes: Optional[int] = None
        self.info_calls = self.input_batch_requests = self.scratch_calls = 0
        self.acquire_calls = self.map_calls = self.seal_calls = self.release_calls = 0
        self.sealed: Optional[OutputSealDescriptor] = None
        self.sealed_metadata: Optional[bytes] = None
        self._views = []
        self._build()

    def _build(self):
        self._batch_views = []
        for b in self.batches:
            arr = (InputItemView * max(1, len(b.items)))()
            for i, it in enumerate(b.items):
                v = arr[i]
                v.abi = header(InputItemView)
                v.payload.abi = header(PayloadView)
                v.payload.data = _address(it.data)
                v.payload.byte_count = _nbytes(it.data) if it.byte_count is None else it.byte_count
                v.payload.memory_domain = it.memory_domain
                v.payload.alignment = it.alignment
                v.payload.type = it.type if it.type is not None else type_descriptor(it.type_ref)
                v.metadata = byte_view(it.metadata)
                v.semantic_key_hash = it.semantic_key_hash
                v.order_key = it.order_key
                v.item_ordinal = it.item_ordinal
            self._batch_views.append((b, arr))
        out_type = self.output_spec.type if self.output_spec.type is not None else type_descriptor(
            self.output_spec.type_ref)
        self._out_type = out_type
        self._grant_token = C.c_uint64(0xA11CE)

        def get_info(ctx, lease, out, err):
            self.info_calls += 1
            i = FiringLeaseInfo()
            i.abi = header(FiringLeaseInfo)
            i.terminal_epoch = self.terminal_epoch
            i.input_batch_count = len(self.batches) if self.input_batch_count is None else self.input_batch_count
            i.output_grant_count = 1
            i.reserved_output_bytes = _nbytes(self.output) if self.reserved_output_bytes is None else \
                self.reserved_output_bytes
            i.reserved_scratch_bytes = self.scratch_bytes if self.reserved_scratch_bytes is None else \
                self.reserved_scratch_bytes
            out[0] = i
            return OK

        def get_input_batch(ctx, lease, index, out, err):
            self.input_batch_requests += 1
            if index >= len(self._batch_views):
                return INVALID_ARGUMENT
            b, arr = self._batch_views[index]
            v = InputBatchView()
            v.abi = header(InputBatchView)
            v.items = C.cast(arr, C.POINTER(InputItemView))
            v.item_count = len(b.items)
            v.input_port = b.port
            v.batch_id = 5 + index
            v.order_domain = 7
            out[0] = v
            return OK

        def get_scratch(ctx, lease, out, err):
            self.scratch_calls += 1
            s = ScratchView()
            s.abi = header(ScratchView)
            s.data = self.scratch.ctypes.data
            s.byte_count = self.scratch_bytes
            s.memory_domain = MEMORY_HOST_PAGEABLE
            s.alignment = 64
            out[0] = s
            return OK

        def acquire(ctx, lease, slot, out_grant, err):
            self.acquire_calls += 1
            out_grant[0] = C.addressof(self._grant_token)
            return OK

        def map_payload(ctx, grant, out, err):
            self.map_calls += 1
            m = MutablePayloadView()
            m.abi = header(MutablePayloadView)
            m.data = _address(self.output)
            m.capacity_bytes = _nbytes(self.output)
            m.committed_bytes = 0
            m.memory_domain = self.output_spec.memory_domain
            m.alignment = self.output_spec.alignment
            m.type = self._out_type
            out[0] = m
            return OK

        def seal(ctx, grant, desc, err):
            self.seal_calls += 1
            d = desc[0]
            self.sealed = OutputSealDescriptor.from_buffer_copy(d)
            self.sealed_metadata = C.string_at(d.metadata.data, d.metadata.size) if d.metadata.data else b""
            return OK

        def release(ctx, grant, err):
            self.release_calls += 1
            return OK

        self._fns = [GetInfoFn(get_info), GetInputBatchFn(get_input_batch), GetScratchFn(get_scratch),
                     AcquireOutputFn(acquire), MapFn(map_payload), SealFn(seal), ReleaseFn(release)]
        self.grants = OutputGrantCallbacks(header(OutputGrantCallbacks), None, self._fns[4], self._fns[5],
                                           self._fns[6])
        cb = FiringLeaseCallbacks()
        cb.abi = header(FiringLeaseCallbacks, self.caps)
        cb.output_grants = C.pointer(self.grants)
        cb.get_info, cb.get_input_batch, cb.get_scratch = self._fns[0], self._fns[1], self._fns[2]
        cb.acquire_output_grant = self._fns[3]
        self.callbacks = cb
        self._lease_token = C.c_uint64(0x1EA5E)

    @property
    def lease_handle(self):
        return C.addressof(self._lease_token)


# ---- module and operators ----

class Operator:
    def __init__(self, module: "ProviderModule", handle: int):
        self.module, self.handle = module, handle
        self.context = self.key_state = None

    def open(self, device_ordinal: int = 0):
        """execution_context_create + key_state_init + operator_on_start; returns the first non-OK status."""
        api = self.module.api
        ctx = vp()
        d = ExecutionContextDescriptor()
        d.abi = header(ExecutionContextDescriptor)
        d.device_ordinal = device_ordinal
        e = empty_error()
        st = api.execution_context_create(self.handle, C.byref(d), C.byref(ctx), C.byref(e))
        if st != OK:
            return st
        self.context = ctx.value
        k = KeyStateDescriptor()
        k.abi = header(KeyStateDescriptor)
        k.semantic_key = byte_view(b"scan")
        ks = vp()
        st = api.key_state_init(self.handle, self.context, C.byref(k), C.byref(ks), C.byref(e))
        if st != OK:
            return st
        self.key_state = ks.value
        s = ScanStartDescriptor()
        s.abi = header(ScanStartDescriptor)
        s.run_id, s.scan_id = utf8("run"), utf8("scan")
        s.normalized_scan_facts_digest.abi = header(Digest256)
        s.execution_plan_digest.abi = header(Digest256)
        return api.operator_on_start(self.handle, self.context, self.key_state, C.byref(s), C.byref(e))

    def process(self, lease: Lease):
        """operator_process_batch -> (status, ProcessResult, error message)."""
        r = ProcessResult()
        r.abi = header(ProcessResult)
        e = empty_error()
        st = self.module.api.operator_process_batch(self.handle, self.context, self.key_state, lease.lease_handle,
                                                    C.byref(lease.callbacks), C.byref(r), C.byref(e))
        return st, r, error_text(e)

    def scan_end(self, lease: Lease, kind: int = SCAN_END_NORMAL, terminal_epoch: int = 9):
        d = ScanEndDescriptor()
        d.abi = header(ScanEndDescriptor)
        d.kind, d.terminal_epoch = kind, terminal_epoch
        r = ProcessResult()
        r.abi = header(ProcessResult)
        e = empty_error()
        st = self.module.api.operator_on_scan_end(self.handle, self.context, self.key_state, C.byref(d),
                                                  lease.lease_handle, C.byref(lease.callbacks), C.byref(r),
                                                  C.byref(e))
        return st, r

    def cancel(self, kind: int = SCAN_END_CANCELLED):
        d = CancelContext()
        d.abi = header(CancelContext)
        d.kind = kind
        d.reason = utf8("stop")
        e = empty_error()
        return self.module.api.operator_on_cancel(self.handle, self.context, self.key_state, C.byref(d), C.byref(e))

    def close(self):
        api = self.module.api
        if self.key_state:
            api.key_state_reset(self.handle, self.context, self.key_state)
            self.key_state = None
        if self.context:
            api.execution_context_destroy(self.handle, self.context)
            self.context = None
        if self.handle:
            api.operator_destroy(self.handle)
            self.handle = None


class ProviderModule:
    """A loaded provider module: libksj_provider_<name>.so beside this file, or in the
    directory KSJ_PROVIDER_DIR names (another build of the same modules, e.g. one with
    the host sanitizers)."""

    def __init__(self, name: str, path: Optional[str] = None):
        self.path = path or os.path.join(os.environ.get("KSJ_PROVIDER_DIR") or PKG, f"libksj_provider_{name}.so")
        if not os.path.exists(self.path):
            raise ImportError(f"{self.path} is not built; run python -m kspacejet_amd.build")
        self.lib = C.CDLL(self.path)
        self.lib.ksj_provider_query.restype = u32
        self.lib.ksj_provider_query.argtypes = [C.POINTER(QueryRequest), C.POINTER(ProviderDescriptor),
                                                C.POINTER(ProviderApi), E]
        req = QueryRequest(header(QueryRequest), 0, utf8("kspacejet_amd.provider_host"))
        self.descriptor = ProviderDescriptor()
        self.descriptor.abi = header(ProviderDescriptor)
        self.api = ProviderApi()
        self.api.abi = header(ProviderApi)
        e = empty_error()
        st = self.lib.ksj_provider_query(C.byref(req), C.byref(self.descriptor), C.byref(self.api), C.byref(e))
        if st != OK:
            raise RuntimeError(f"ksj_provider_query failed ({st}): {error_text(e)}")

    @property
    def provider_id(self) -> str:
        return view_text(self.descriptor.provider_id)

    @property
    def bundle_digest(self) -> bytes:
        return bytes(self.descriptor.bundle_digest.bytes)

    @property
    def operators(self) -> List[OperatorDescriptor]:
        return [self.descriptor.operators[i] for i in range(self.descriptor.operator_count)]

    def create(self, operator_id: str, config):
        """operator_create -> (status, Operator or None, error message); config is text
        (sent as UTF-8) or raw bytes."""
        req = OperatorCreateRequest()
        req.abi = header(OperatorCreateRequest)
        req.operator_id = utf8(operator_id)
        raw = config if isinstance(config, (bytes, bytearray)) else config.encode()
        req.canonical_config = byte_view(bytes(raw))
        out = vp()
        e = empty_error()
        st = self.api.operator_create(C.byref(req), C.byref(out), C.byref(e))
        return st, (Operator(self, out.value) if st == OK else None), error_text(e)
