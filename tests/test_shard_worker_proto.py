"""The shard RPC messages (vecsearch_amd.search_proto) against the reference's search.proto: the
bytes each message makes are the proto3 encodings of its fields (pinned by hand below), and,
where the reference tree is present, every field's name, number and type match the .proto text."""
import os
import re

import pytest

from vecsearch_amd import search_proto as sp

REF_PROTO = os.path.join(os.environ.get("VS_REFERENCE", "/root/reference"), "distributed", "search.proto")


def test_wire_bytes():
    # field 1 bytes (tag 0x0a, length 4), field 2 varint 10 (0x10 0x0a), field 3 varint 64 (0x18 0x40)
    r = sp.SearchRequest(query=b"\x00\x00\x80\x3f", k=10, ef=64)
    assert r.SerializeToString() == bytes.fromhex("0a040000803f100a1840")
    # packed int64 [5, 300] (0x0a, len 3, 05 ac 02), packed float [0.5, 1.0] (0x12, len 8), one map
    # entry {"shard_0": 1.5} as a nested message (0x1a, len 18: key 0x0a, value 0x11 double)
    s = sp.SearchResponse(ids=[5, 300], distances=[0.5, 1.0], timings_ms={"shard_0": 1.5})
    assert s.SerializeToString() == bytes.fromhex("0a0305ac0212080000003f0000803f1a120a0773686172645f3011000000000000f83f")
    assert sp.InfoResponse(num_vectors=1_000_000, dim=384, num_shards=1).SerializeToString() == \
        bytes.fromhex("08c0843d1080031801")
    assert sp.InfoRequest().SerializeToString() == b""
    back = sp.SearchResponse.FromString(s.SerializeToString())
    assert list(back.ids) == [5, 300] and dict(back.timings_ms) == {"shard_0": 1.5}
    assert sp.SEARCH_METHOD == "/vecsearch.Searcher/Search" and sp.INFO_METHOD == "/vecsearch.Searcher/Info"


@pytest.mark.skipif(not os.path.exists(REF_PROTO), reason="the reference sources are not present")
def test_fields_match_reference_proto():
    text = open(REF_PROTO).read()
    assert re.search(r"^package\s+vecsearch;", text, re.M)
    assert re.search(r"rpc\s+Search\(SearchRequest\)\s+returns\s+\(SearchResponse\)", text)
    assert re.search(r"rpc\s+Info\(InfoRequest\)\s+returns\s+\(InfoResponse\)", text)
    types = {"bytes": 12, "int32": 5, "int64": 3, "float": 2, "double": 1, "string": 9}
    for m in re.finditer(r"message\s+(\w+)\s*\{([^}]*)\}", text):
        desc = sp.FILE.message_types_by_name[m.group(1)]
        declared = re.findall(r"(repeated\s+)?(map<\w+,\s*\w+>|\w+)\s+(\w+)\s*=\s*(\d+);", m.group(2))
        assert len(declared) == len(desc.fields)
        for rep, typ, name, num in declared:
            f = desc.fields_by_name[name]
            assert f.number == int(num)
            if typ.startswith("map<"):
                assert f.message_type.GetOptions().map_entry
            else:
                assert f.type == types[typ] and f.is_repeated == bool(rep)
