"""The messages and service of the reference's shard RPC (distributed/search.proto:1-28),
declared through the protobuf descriptor API (no generated code, no protoc needed).

Same package (vecsearch), service (Searcher: Search, Info), message names, field numbers and
types as search.proto, so the bytes on the wire are those of the reference's generated
search_pb2 and its coordinator (distributed/coordinator.py) and clients talk to a worker built
on these classes unchanged:

    SearchRequest   bytes query = 1 (little-endian float32[dim]); int32 k = 2; int32 ef = 3
    SearchResponse  repeated int64 ids = 1; repeated float distances = 2;
                    map<string, double> timings_ms = 3
    InfoRequest     (empty)
    InfoResponse    int64 num_vectors = 1; int32 dim = 2; int32 num_shards = 3
"""
from google.protobuf import descriptor_pb2, descriptor_pool, message_factory

PACKAGE = "vecsearch"
SERVICE = "vecsearch.Searcher"

_F = descriptor_pb2.FieldDescriptorProto


def _file_proto():
    fp = descriptor_pb2.FileDescriptorProto(name="vecsearch_amd/search.proto", package=PACKAGE, syntax="proto3")

    def message(name, fields):
        m = fp.message_type.add(name=name)
        for fname, number, ftype, label, type_name in fields:
            f = m.field.add(name=fname, number=number, type=ftype, label=label)
            if type_name:
                f.type_name = type_name
        return m

    opt, rep = _F.LABEL_OPTIONAL, _F.LABEL_REPEATED
    message("SearchRequest", [("query", 1, _F.TYPE_BYTES, opt, None), ("k", 2, _F.TYPE_INT32, opt, None),
                              ("ef", 3, _F.TYPE_INT32, opt, None)])
    resp = message("SearchResponse", [("ids", 1, _F.TYPE_INT64, rep, None),
                                      ("distances", 2, _F.TYPE_FLOAT, rep, None),
                                      ("timings_ms", 3, _F.TYPE_MESSAGE, rep, ".vecsearch.SearchResponse.TimingsMsEntry")])
    entry = resp.nested_type.add(name="TimingsMsEntry")  # what protoc makes of map<string, double>
    entry.field.add(name="key", number=1, type=_F.TYPE_STRING, label=opt)
    entry.field.add(name="value", number=2, type=_F.TYPE_DOUBLE, label=opt)
    entry.options.map_entry = True
    message("InfoRequest", [])
    message("InfoResponse", [("num_vectors", 1, _F.TYPE_INT64, opt, None), ("dim", 2, _F.TYPE_INT32, opt, None),
                             ("num_shards", 3, _F.TYPE_INT32, opt, None)])
    svc = fp.service.add(name="Searcher")
    svc.method.add(name="Search", input_type=".vecsearch.SearchRequest", output_type=".vecsearch.SearchResponse")
    svc.method.add(name="Info", input_type=".vecsearch.InfoRequest", output_type=".vecsearch.InfoResponse")
    return fp


_pool = descriptor_pool.DescriptorPool()
_file = _pool.Add(_file_proto())
FILE = _pool.FindFileByName("vecsearch_amd/search.proto")

SearchRequest = message_factory.GetMessageClass(_pool.FindMessageTypeByName("vecsearch.SearchRequest"))
SearchResponse = message_factory.GetMessageClass(_pool.FindMessageTypeByName("vecsearch.SearchResponse"))
InfoRequest = message_factory.GetMessageClass(_pool.FindMessageTypeByName("vecsearch.InfoRequest"))
InfoResponse = message_factory.GetMessageClass(_pool.FindMessageTypeByName("vecsearch.InfoResponse"))

# gRPC method paths, as a generated search_pb2_grpc names them
SEARCH_METHOD = f"/{SERVICE}/Search"
INFO_METHOD = f"/{SERVICE}/Info"


def add_searcher_to_server(servicer, server):
    """Registers servicer.Search / servicer.Info (callables (request, context) -> response) as the
    vecsearch.Searcher service, the role of search_pb2_grpc.add_SearcherServicer_to_server."""
    import grpc

    handlers = {
        "Search": grpc.unary_unary_rpc_method_handler(servicer.Search, request_deserializer=SearchRequest.FromString,
                                                      response_serializer=SearchResponse.SerializeToString),
        "Info": grpc.unary_unary_rpc_method_handler(servicer.Info, request_deserializer=InfoRequest.FromString,
                                                    response_serializer=InfoResponse.SerializeToString),
    }
    server.add_generic_rpc_handlers((grpc.method_handlers_generic_handler(SERVICE, handlers),))


class SearcherStub:
    """Client of the service (the role of search_pb2_grpc.SearcherStub) over a grpc channel."""

    def __init__(self, channel):
        self.Search = channel.unary_unary(SEARCH_METHOD, request_serializer=SearchRequest.SerializeToString,
                                          response_deserializer=SearchResponse.FromString)
        self.Info = channel.unary_unary(INFO_METHOD, request_serializer=InfoRequest.SerializeToString,
                                        response_deserializer=InfoResponse.FromString)
