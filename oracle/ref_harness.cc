// ref_harness.cc — drives the reference's own CPU table and compaction code,
// compiled from its sources under /root/reference by oracle/Makefile into
// oracle/_ref/libluda_ref.so.  TEST INFRASTRUCTURE ONLY: used to pin the C
// restatement (oracle/lsm_oracle.c) and, where present, as the CPU baseline.
//
// What runs is the reference's code: TableBuilder/BlockBuilder/FilterBlock
// (table/), the bloom policy and crc32c (util/), InternalKeyComparator and
// InternalFilterPolicy (db/dbformat.cc), Table::Open and its iterators,
// NewMergingIterator and NewTwoLevelIterator (table/), and
// Compaction::ShouldStopBefore / IsBaseLevelForKey (db/version_set.cc), whose
// members this fork made public (db/version_set.h:362).  Only the driving loop
// of DBImpl::DoCompactionWork is restated here, because db/db_impl.cc cannot
// be built without CUDA (db/db_impl.cc:43-44): the loop below follows
// db/db_impl.cc:1090-1197 statement by statement.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "db/dbformat.h"
#include "db/log_writer.h"
#include "db/table_cache.h"
#include "db/version_edit.h"
#include "db/version_set.h"
#include "port/port.h"
#include "leveldb/comparator.h"
#include "leveldb/env.h"
#include "leveldb/filter_policy.h"
#include "leveldb/iterator.h"
#include "leveldb/options.h"
#include "leveldb/table.h"
#include "leveldb/table_builder.h"
#include "table/merger.h"
#include "table/two_level_iterator.h"

#include "../include/luda.h"

namespace {

using leveldb::Slice;
using leveldb::Status;

class StringSink : public leveldb::WritableFile {
 public:
  std::string contents;
  Status Append(const Slice& d) override {
    contents.append(d.data(), d.size());
    return Status::OK();
  }
  Status Close() override { return Status::OK(); }
  Status Flush() override { return Status::OK(); }
  Status Sync() override { return Status::OK(); }
};

class MemSource : public leveldb::RandomAccessFile {
 public:
  MemSource(const uint8_t* p, uint64_t n) : p_(p), n_(n) {}
  Status Read(uint64_t offset, size_t n, Slice* result, char* scratch) const override {
    if (offset > n_) return Status::InvalidArgument("read past end");
    if (offset + n > n_) n = n_ - offset;
    memcpy(scratch, p_ + offset, n);
    *result = Slice(scratch, n);
    return Status::OK();
  }

 private:
  const uint8_t* p_;
  uint64_t n_;
};

// Results in the engine's luda_result layout.
struct Results {
  std::vector<luda_output> outs;
  std::string keys;
  std::string data;
  void Add(const std::string& table, uint64_t nent, const std::string& smallest, const std::string& largest) {
    luda_output o;
    memset(&o, 0, sizeof(o));
    o.offset = data.size();
    o.file_size = table.size();
    o.num_entries = nent;
    o.smallest_offset = keys.size();
    o.smallest_size = (uint32_t)smallest.size();
    keys += smallest;
    o.largest_offset = keys.size();
    o.largest_size = (uint32_t)largest.size();
    keys += largest;
    data += table;
    outs.push_back(o);
  }
  luda_result* Finish(uint64_t bytes_read) {
    luda_result* r = (luda_result*)calloc(1, sizeof(luda_result));
    r->num_outputs = (uint32_t)outs.size();
    r->outputs = (luda_output*)malloc(sizeof(luda_output) * (outs.size() ? outs.size() : 1));
    if (!outs.empty()) memcpy(r->outputs, outs.data(), sizeof(luda_output) * outs.size());
    r->keys = (uint8_t*)malloc(keys.size() ? keys.size() : 1);
    memcpy(r->keys, keys.data(), keys.size());
    r->data = (uint8_t*)malloc(data.size() ? data.size() : 1);
    memcpy(r->data, data.data(), data.size());
    r->total_bytes = data.size();
    r->stats.bytes_read = bytes_read;
    r->stats.bytes_written = data.size();
    r->stats.attempts = 1;
    return r;
  }
};

// The options DBImpl hands to TableBuilder after SanitizeOptions
// (db/db_impl.cc:235-260): internal comparator and internal filter policy.
struct RefOptions {
  leveldb::InternalKeyComparator icmp;
  const leveldb::FilterPolicy* bloom;
  leveldb::InternalFilterPolicy* ifp;
  leveldb::Options opt;
  explicit RefOptions(const luda_options& o)
      : icmp(leveldb::BytewiseComparator()), bloom(nullptr), ifp(nullptr) {
    opt.comparator = &icmp;
    opt.block_size = o.block_size;
    opt.block_restart_interval = (int)o.block_restart_interval;
    opt.max_file_size = o.max_file_size;
    opt.compression = leveldb::kNoCompression;
    if (o.bloom_bits_per_key) {
      bloom = leveldb::NewBloomFilterPolicy((int)o.bloom_bits_per_key);
      ifp = new leveldb::InternalFilterPolicy(bloom);
      opt.filter_policy = ifp;
    }
  }
  ~RefOptions() {
    delete ifp;
    delete bloom;
  }
};

// Index iterator over a level's file list for NewTwoLevelIterator, standing
// in for Version::LevelFileNumIterator (db/version_set.cc:166-218), which is
// private to version_set.cc.  Compaction only calls SeekToFirst and Next.
class FileListIterator : public leveldb::Iterator {
 public:
  explicit FileListIterator(size_t n) : n_(n), i_(n) {}
  bool Valid() const override { return i_ < n_; }
  void SeekToFirst() override { i_ = 0; }
  void SeekToLast() override { i_ = n_ ? n_ - 1 : n_; }
  void Seek(const Slice&) override { i_ = 0; }
  void Next() override { i_++; }
  void Prev() override { i_ = i_ == 0 ? n_ : i_ - 1; }
  Slice key() const override { return Slice(); }
  Slice value() const override {
    EncodeIndex();
    return Slice(buf_, 8);
  }
  Status status() const override { return Status::OK(); }

 private:
  void EncodeIndex() const {
    uint64_t v = i_;
    memcpy(buf_, &v, 8);
  }
  size_t n_, i_;
  mutable char buf_[8];
};

struct LevelArg {
  std::vector<leveldb::Table*>* tables;
};

leveldb::Iterator* OpenFileIter(void* arg, const leveldb::ReadOptions& ro, const Slice& v) {
  LevelArg* la = reinterpret_cast<LevelArg*>(arg);
  uint64_t i;
  memcpy(&i, v.data(), 8);
  return (*la->tables)[i]->NewIterator(ro);
}

}  // namespace

extern "C" {

void ref_free_result(luda_result* r) {
  if (!r) return;
  free(r->outputs);
  free(r->keys);
  free(r->data);
  free(r);
}

// BuildTable (db/builder.cc:17-77) plus the DoCompactionWork size cut.
int ref_build_tables(const luda_options* o, uint64_t n, const uint8_t* keys, const uint64_t* key_off,
                     const uint8_t* vals, const uint64_t* val_off, luda_result** out) {
  RefOptions ro(*o);
  Results res;
  StringSink* sink = nullptr;
  leveldb::TableBuilder* b = nullptr;
  std::string smallest, largest;
  for (uint64_t i = 0; i < n; i++) {
    Slice k((const char*)keys + key_off[i], key_off[i + 1] - key_off[i]);
    Slice v((const char*)vals + val_off[i], val_off[i + 1] - val_off[i]);
    if (!b) {
      sink = new StringSink;
      b = new leveldb::TableBuilder(ro.opt, sink);
    }
    if (b->NumEntries() == 0) smallest = k.ToString();
    largest = k.ToString();
    b->Add(k, v);
    if (b->FileSize() >= o->max_file_size) {
      b->Finish();
      res.Add(sink->contents, b->NumEntries(), smallest, largest);
      delete b;
      delete sink;
      b = nullptr;
    }
  }
  if (b) {
    b->Finish();
    res.Add(sink->contents, b->NumEntries(), smallest, largest);
    delete b;
    delete sink;
  }
  *out = res.Finish(0);
  return LUDA_OK;
}

int ref_compact(const luda_job* job, luda_result** out) {
  *out = nullptr;
  RefOptions ro(job->options);
  leveldb::Options& opt = ro.opt;
  leveldb::ReadOptions rdo;
  rdo.verify_checksums = false;  // the fork never checks (table/format.cc:88)
  rdo.fill_cache = false;

  // Open every input table.
  std::vector<MemSource*> sources;
  std::vector<leveldb::Table*> tables[2];
  uint64_t bytes_read = 0;
  for (int which = 0; which < 2; which++) {
    for (uint32_t i = 0; i < job->num_inputs[which]; i++) {
      const luda_file& f = job->inputs[which][i];
      MemSource* src = new MemSource(f.data, f.size);
      sources.push_back(src);
      leveldb::Table* t = nullptr;
      Status s = leveldb::Table::Open(opt, src, f.size, &t);
      if (!s.ok()) {
        for (auto* x : sources) delete x;
        for (auto& v : tables)
          for (auto* x : v) delete x;
        return LUDA_CORRUPTION;
      }
      tables[which].push_back(t);
      bytes_read += f.size;
    }
  }

  // VersionSet::MakeInputIterator, db/version_set.cc:1237-1268.
  LevelArg largs[2] = {{&tables[0]}, {&tables[1]}};
  std::vector<leveldb::Iterator*> list;
  for (int which = 0; which < 2; which++) {
    if (tables[which].empty()) continue;
    if (job->level + which == 0) {
      for (auto* t : tables[which]) list.push_back(t->NewIterator(rdo));
    } else {
      list.push_back(leveldb::NewTwoLevelIterator(new FileListIterator(tables[which].size()), &OpenFileIter,
                                                  &largs[which], rdo));
    }
  }
  leveldb::Iterator* input = leveldb::NewMergingIterator(&ro.icmp, list.data(), (int)list.size());

  // The Compaction and the Version it reads, populated from the job.
  leveldb::TableCache tc("refdb", opt, 16);
  leveldb::VersionSet vset("refdb", &opt, &tc, &ro.icmp);
  leveldb::Version* v = new leveldb::Version(&vset);
  for (uint32_t i = 0; i < job->num_deeper; i++) {
    const luda_level_file& lf = job->deeper[i];
    if (lf.level < 0 || lf.level >= leveldb::config::kNumLevels) continue;
    leveldb::FileMetaData* f = new leveldb::FileMetaData;
    f->refs = 1;
    std::string s((const char*)lf.smallest_user_key.data, lf.smallest_user_key.size);
    std::string l((const char*)lf.largest_user_key.data, lf.largest_user_key.size);
    s.append(8, '\0');
    l.append(8, '\0');
    f->smallest.DecodeFrom(s);
    f->largest.DecodeFrom(l);
    v->files_[lf.level].push_back(f);
  }
  leveldb::Compaction* c = new leveldb::Compaction(&opt, job->level);
  c->input_version_ = v;
  v->Ref();
  std::vector<leveldb::FileMetaData*> gps;
  for (uint32_t i = 0; i < job->num_grandparents; i++) {
    leveldb::FileMetaData* f = new leveldb::FileMetaData;
    f->largest.DecodeFrom(Slice((const char*)job->grandparents[i].largest.data, job->grandparents[i].largest.size));
    f->file_size = job->grandparents[i].file_size;
    gps.push_back(f);
  }
  c->grandparents_ = gps;

  // DBImpl::DoCompactionWork loop, db/db_impl.cc:1090-1197.
  Results res;
  StringSink* outfile = nullptr;
  leveldb::TableBuilder* builder = nullptr;
  std::string smallest, largest;
  leveldb::ParsedInternalKey ikey;
  std::string current_user_key;
  bool has_current_user_key = false;
  leveldb::SequenceNumber last_sequence_for_key = leveldb::kMaxSequenceNumber;
  const leveldb::SequenceNumber smallest_snapshot = job->smallest_snapshot;
  const leveldb::Comparator* ucmp = leveldb::BytewiseComparator();
  uint64_t total = 0, kept = 0;
  auto finish_output = [&]() {
    builder->Finish();
    res.Add(outfile->contents, builder->NumEntries(), smallest, largest);
    delete builder;
    delete outfile;
    builder = nullptr;
    outfile = nullptr;
  };
  for (input->SeekToFirst(); input->Valid();) {
    Slice key = input->key();
    total++;
    if (c->ShouldStopBefore(key) && builder != nullptr) finish_output();
    bool drop = false;
    if (!leveldb::ParseInternalKey(key, &ikey)) {
      current_user_key.clear();
      has_current_user_key = false;
      last_sequence_for_key = leveldb::kMaxSequenceNumber;
    } else {
      if (!has_current_user_key || ucmp->Compare(ikey.user_key, Slice(current_user_key)) != 0) {
        current_user_key.assign(ikey.user_key.data(), ikey.user_key.size());
        has_current_user_key = true;
        last_sequence_for_key = leveldb::kMaxSequenceNumber;
      }
      if (last_sequence_for_key <= smallest_snapshot) {
        drop = true;
      } else if (ikey.type == leveldb::kTypeDeletion && ikey.sequence <= smallest_snapshot &&
                 c->IsBaseLevelForKey(ikey.user_key)) {
        drop = true;
      }
      last_sequence_for_key = ikey.sequence;
    }
    if (!drop) {
      if (builder == nullptr) {
        outfile = new StringSink;
        builder = new leveldb::TableBuilder(opt, outfile);
      }
      if (builder->NumEntries() == 0) smallest = key.ToString();
      largest = key.ToString();
      builder->Add(key, input->value());
      kept++;
      if (builder->FileSize() >= c->MaxOutputFileSize()) finish_output();
    }
    input->Next();
  }
  if (builder != nullptr) finish_output();
  Status st = input->status();
  delete input;
  delete c;  // unrefs v, which drops the deeper-level files
  for (auto* f : gps) delete f;
  for (auto& vv : tables)
    for (auto* t : vv) delete t;
  for (auto* s : sources) delete s;
  if (!st.ok()) return LUDA_CORRUPTION;
  *out = res.Finish(bytes_read);
  (*out)->stats.entries_in = total;
  (*out)->stats.entries_kept = kept;
  return LUDA_OK;
}

// VersionEdit::EncodeTo (db/version_edit.cc:41-89) of an edit described by
// a luda_edit, built with VersionEdit's own setters.
int ref_encode_edit(const luda_edit* e, uint8_t* dst, uint64_t cap, uint64_t* len) {
  leveldb::VersionEdit ve;
  if (e->has & LUDA_EDIT_HAS_COMPARATOR)
    ve.SetComparatorName(Slice((const char*)e->comparator.data, e->comparator.size));
  if (e->has & LUDA_EDIT_HAS_LOG_NUMBER) ve.SetLogNumber(e->log_number);
  if (e->has & LUDA_EDIT_HAS_PREV_LOG_NUMBER) ve.SetPrevLogNumber(e->prev_log_number);
  if (e->has & LUDA_EDIT_HAS_NEXT_FILE) ve.SetNextFile(e->next_file_number);
  if (e->has & LUDA_EDIT_HAS_LAST_SEQUENCE) ve.SetLastSequence(e->last_sequence);
  for (uint32_t i = 0; i < e->num_compact_pointers; i++) {
    leveldb::InternalKey k;
    k.DecodeFrom(Slice((const char*)e->compact_pointers[i].key.data, e->compact_pointers[i].key.size));
    ve.SetCompactPointer(e->compact_pointers[i].level, k);
  }
  for (uint32_t i = 0; i < e->num_deleted; i++) ve.DeleteFile(e->deleted[i].level, e->deleted[i].number);
  for (uint32_t i = 0; i < e->num_new_files; i++) {
    const luda_edit_file& f = e->new_files[i];
    leveldb::InternalKey a, b;
    a.DecodeFrom(Slice((const char*)f.smallest.data, f.smallest.size));
    b.DecodeFrom(Slice((const char*)f.largest.data, f.largest.size));
    ve.AddFile(f.level, f.number, f.file_size, a, b);
  }
  std::string out;
  ve.EncodeTo(&out);
  *len = out.size();
  if (out.size() > cap) return LUDA_INVALID_ARGUMENT;
  memcpy(dst, out.data(), out.size());
  return LUDA_OK;
}

// VersionSet::PickCompaction (db/version_set.cc:1270-1464) over a tree given
// as an edit (its new files and compact pointers), applied to an empty
// VersionSet in directory `dbdir` with LogAndApply (which also runs Finalize,
// :1050-1086).  Out: the picked level (-1 when none), the file numbers of
// inputs_[0], inputs_[1] and grandparents_ in the Compaction's order,
// IsTrivialMove, and the Compaction's own edit (its compact pointer) encoded.
int ref_pick(const char* dbdir, const luda_edit* tree, uint64_t max_file_size, int* level, uint64_t* in0,
             uint32_t* n0, uint64_t* in1, uint32_t* n1, uint64_t* gp, uint32_t* ngp, int* trivial, uint8_t* edit,
             uint64_t edit_cap, uint64_t* edit_len) {
  leveldb::Options options;
  options.max_file_size = max_file_size;
  leveldb::Env* env = leveldb::Env::Default();
  options.env = env;
  env->CreateDir(dbdir);
  leveldb::InternalKeyComparator icmp(leveldb::BytewiseComparator());
  leveldb::TableCache tc(dbdir, options, 10);
  leveldb::VersionSet vs(dbdir, &options, &tc, &icmp);
  leveldb::VersionEdit ve;
  for (uint32_t i = 0; i < tree->num_compact_pointers; i++) {
    leveldb::InternalKey k;
    k.DecodeFrom(Slice((const char*)tree->compact_pointers[i].key.data, tree->compact_pointers[i].key.size));
    ve.SetCompactPointer(tree->compact_pointers[i].level, k);
  }
  for (uint32_t i = 0; i < tree->num_new_files; i++) {
    const luda_edit_file& f = tree->new_files[i];
    leveldb::InternalKey a, b;
    a.DecodeFrom(Slice((const char*)f.smallest.data, f.smallest.size));
    b.DecodeFrom(Slice((const char*)f.largest.data, f.largest.size));
    ve.AddFile(f.level, f.number, f.file_size, a, b);
  }
  leveldb::port::Mutex mu;
  mu.Lock();
  Status st = vs.LogAndApply(&ve, &mu);
  mu.Unlock();
  if (!st.ok()) return LUDA_IO_ERROR;
  leveldb::Compaction* c = vs.PickCompaction();
  *level = -1;
  *n0 = *n1 = *ngp = 0;
  *trivial = 0;
  *edit_len = 0;
  if (!c) return LUDA_OK;
  *level = c->level();
  for (auto* f : c->inputs_[0]) in0[(*n0)++] = f->number;
  for (auto* f : c->inputs_[1]) in1[(*n1)++] = f->number;
  for (auto* f : c->grandparents_) gp[(*ngp)++] = f->number;
  *trivial = c->IsTrivialMove() ? 1 : 0;
  std::string enc;
  c->edit()->EncodeTo(&enc);
  *edit_len = enc.size();
  if (enc.size() <= edit_cap) memcpy(edit, enc.data(), enc.size());
  c->ReleaseInputs();
  delete c;
  return LUDA_OK;
}

// log::Writer::AddRecord (db/log_writer.cc:35-80) of one record appended to a
// log of length *block_offset (only its offset in the block matters).
int ref_log_record(uint64_t* block_offset, const uint8_t* payload, uint64_t n, uint8_t* dst, uint64_t cap,
                   uint64_t* len) {
  StringSink sink;
  {
    leveldb::log::Writer w(&sink, *block_offset);
    w.AddRecord(Slice((const char*)payload, n));
  }
  *len = sink.contents.size();
  if (sink.contents.size() > cap) return LUDA_INVALID_ARGUMENT;
  memcpy(dst, sink.contents.data(), sink.contents.size());
  *block_offset = (*block_offset + sink.contents.size()) % 32768;
  return LUDA_OK;
}

}  // extern "C"
