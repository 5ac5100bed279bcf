c **********************************************************************
c * ref_harness: test-only driver that links the reference's own      *
c * Fortran objects (compiled from mef/src by oracle/Makefile.ref)     *
c * and runs the path of the `mesh` + `solvm` macros of Main.f         *
c * (Main.f:348-534, 718-849) on one deck, then writes the arrays the  *
c * oracle and the HIP path are checked against.                       *
c *                                                                    *
c * It is the checker's checker: nothing in the product calls it.      *
c *                                                                    *
c * usage: ref_harness deck outdir solvtol maxit nth mode [maxmem]     *
c *   mode 1 = PCG through solv (Solv.f:13), nth>1 -> pcg_omp          *
c *   mode 2 = + gmres(k=50) called with its own signature (Iters.f:528)*
c *   mode 4 = + pbicgstab with its own signature (Iters.f:1079)       *
c *   mode 8 = + block_precond nblock=3 (Precond.f:163) and one      *
c *            application op_block_precond (Precond.f:419)            *
c *   mode 16 = element loop by colours (omp_elmt, Pform.f:108-184)    *
c *   mode 32 = + nodal stresses (tform_mec, Pform.f:258) and the      *
c *             result files of write_mesh_res_mec (Write_vtk.f:742):  *
c *             legacy ASCII, legacy binary and XML ASCII              *
c *   mode 64 = set-up only: stop after the set-up dumps and the     *
c *             product (no solve; for the full-size workloads)        *
c *   mode 256 = + block_precond nblock=2 (2x2 inverses of equation    *
c *             pairs, Precond.f:163-358) and op_block_precond         *
c *   mode 128 = solve only: `deck` is a directory holding a CSRC      *
c *             system (sys.txt: neq nad; ia.i4 ja.i4 ad.f8 al.f8      *
c *             b.f8), solved by solv (Solv.f:13) with nth threads;    *
c *             writes meta.txt (neq nad nth / solv time), solver.log  *
c *             and x.f8 (the reference's CPU baseline on a system     *
c *             assembled elsewhere)                                   *
c *   (modes add up)                                                   *
c * Output (raw little-endian stream files in outdir):                 *
c *   meta.txt, inum.i4, id.i4, ia.i4, ja.i4, ad.f8, al.f8, b.f8,      *
c *   elcolor.i4, colorg.i4 (coloredmesh, Colormsh.f:9-63),            *
c *   x.f8, spmv_v.f8, spmv_y.f8, resid.f8, and optionally             *
c *   x_gmres.f8, x_bicgstab.f8, bdiag.f8, bdiag_y.f8, bdiag2.f8,      *
c *   bdiag2_y.f8, tx.f8, u.f8,                                        *
c *   and the files                                                    *
c *   res_istep_0.vtk, resb_istep_0.vtk, res_istep_0.vtu               *
c **********************************************************************
      program ref_harness
      use Malloc
      implicit none
      include 'mpif.h'
      include 'string.fi'
      include 'transiente.fi'
      include 'parallel.fi'
      include 'gravity.fi'
      include 'time.fi'
      include 'openmp.fi'
      integer nnodev,nnode,numel,numat,nen,nenv,ndf,ndm,nst
      integer neq,nad,maxit,nth,mode,i,nin,nlit,ngram,numcolors
      integer neqf1i,neqf2i
      integer*8 i_ix,i_id,i_ie,i_nload,i_eload,i_e,i_x,i_inum
      integer*8 i_f,i_u,i_tx0,i_fnno
      integer*8 i_xl,i_ul,i_pl,i_sl,i_ld,i_txnl
      integer*8 i_ia,i_ja,i_ad,i_au,i_al,i_b,i_b0,i_x0,i_m,i_bst0
      integer*8 i_z,i_r,i_s,i_v,i_y,i_g,i_h,i_c,i_t,i_p,i_r0
      integer*8 i_colorg,i_elcolor
      integer*8 i_bd,i_aux,mm
      integer param(4),nbd,ntn
      integer*8 i_tx,i_ic
      logical pflag(10)
      character*80 prename
      real*8 solvtol,tt,dot_par
      logical fstress0,fmec,pq
      character*8 sia,sja,sau,sal,sad,mc
      character*256 deck,outd,arg
      external matvec_csrcsym1,dot_par
c ......................................................................
      call MPI_INIT(ierr)
      call MPI_COMM_SIZE(MPI_COMM_WORLD,nprcs,ierr)
      call MPI_COMM_RANK(MPI_COMM_WORLD,my_id,ierr)
      mpi = .false.
      ovlp = .false.
      novlp = .false.
      t = 0.d0
      dt = 1.d0
      alfa = 1.d0
      beta = 1.d0
      gravity(1) = 0.d0
      gravity(2) = 0.d0
      gravity(3) = -9.81d0
      gravity_mod = 9.81d0
      fstress0 = .false.
      pq = .false.
      neqf1i = 0
      neqf2i = 0
c ... arguments
      call get_command_argument(1,deck)
      call get_command_argument(2,outd)
      call get_command_argument(3,arg)
      read(arg,*) solvtol
      call get_command_argument(4,arg)
      read(arg,*) maxit
      call get_command_argument(5,arg)
      read(arg,*) nth
      call get_command_argument(6,arg)
      read(arg,*) mode
      if (command_argument_count() .ge. 7) then
        call get_command_argument(7,arg)
        read(arg,*) maxmem
      endif
      omp_elmt = iand(mode,16) .ne. 0
      nth_elmt = 1
      if (omp_elmt) nth_elmt = 2
      omp_solv = nth .gt. 1
      nth_solv = max(nth,1)
c ... memory pool and deck (Main.f:300-330)
      call init_malloc()
      if (iand(mode,128) .ne. 0) then
c ... solve only, timed as soltime is (Main.f:819-827)
        open(21,file=trim(deck)//'/sys.txt',status='old')
        read(21,*) neq,nad
        close(21)
        i_ia = alloc_4('ia      ',1,neq+1)
        i_ja = alloc_4('ja      ',1,nad)
        i_ad = alloc_8('ad      ',1,neq)
        i_al = alloc_8('al      ',1,nad)
        i_b  = alloc_8('b       ',1,neq)
        i_m  = alloc_8('m       ',1,neq)
        i_x0 = alloc_8('x0      ',1,neq)
        i_fmap = alloc_4('fmap    ',1,1)
        i_xf   = alloc_8('xf      ',1,1)
        i_rcvs = alloc_4('rcvs    ',1,1)
        i_dspl = alloc_4('dspl    ',1,1)
        call ri4(trim(deck)//'/ia.i4',ia(i_ia),neq+1)
        call ri4(trim(deck)//'/ja.i4',ia(i_ja),nad)
        call rf8(trim(deck)//'/ad.f8',ia(i_ad),neq)
        call rf8(trim(deck)//'/al.f8',ia(i_al),nad)
        call rf8(trim(deck)//'/b.f8',ia(i_b),neq)
        call azero(ia(i_x0),neq)
        neqf1 = 0
        neqf2 = 0
        neq3 = 0
        neq4 = 0
        neq_dot = neq
        open(10,file=trim(outd)//'/solver.log')
        tt = MPI_Wtime()
        call solv(neq,nad,ia(i_ia),ia(i_ja),ia(i_ad),ia(i_al),
     .            ia(i_al),ia(i_m),ia(i_b),ia(i_x0),solvtol,maxit,50,
     .            .false.,1,neqf1,neqf2,neq3,neq4,neq_dot,i_fmap,i_xf,
     .            i_rcvs,i_dspl)
        tt = MPI_Wtime() - tt
        open(20,file=trim(outd)//'/meta.txt')
        write(20,*) neq,nad,nth
        write(20,*) tt
        call wf8(trim(outd)//'/x.f8',ia(i_x0),neq)
        goto 900
      endif
      nin = 1
      open(nin,file=deck,status='old',action='read')
   10 continue
      call readmacro(nin,.true.)
      write(mc,'(8a)') (word(i),i=1,8)
      if (mc .ne. 'mesh    ') goto 10
      call rdat(nnode,nnodev,numel,numat,nen,nenv,ndf,ndm,nst,i_ix,
     .          i_ie,i_inum,i_e,i_x,i_id,i_nload,i_eload,i_f,
     .          i_u,i_tx0,i_fnno,fstress0,fmec,pq,nin)
      close(nin)
c ... RCM (Main.f:420-424) and equation numbers (Main.f:426)
      call reord(ia(i_ix),ia(i_inum),nnode,nnode,numel,nen,.true.)
      call numeq(ia(i_id),ia(i_inum),ia(i_id),nnode,ndf,neq)
      call init_front(i_noLG,i_noGL,nno1,nno2,nno3,nnofi,
     .                nno_pload,nnovG,nnoG,nelG,nnodev,nnode,numel,
     .                ovlp,novlp,nprcs,nviz1,nviz2,i_rreqs,i_sreqs,
     .                i_rcvs0i,i_dspl0i,i_fmap0i,mpi)
      call frontb(ndf,ia(i_id),neq,neq1,neq2,neq3,neq4,neq1a,neqf1,
     .            neqf2,neq32,neq_dot,i_fmap,i_rcvs,i_dspl,i_xf,
     .            'fmap    ','rcvs    ','dspl    ','xf      ',0)
      neqf1 = 0
      neqf2 = 0
c ... element work arrays (Main.f:464-473)
      i_xl   = alloc_8('xl      ',ndm,nenv)
      i_ul   = alloc_8('ul      ',1  ,nst)
      i_pl   = alloc_8('pl      ',1  ,nst)
      i_txnl = alloc_8('txnl    ',  6,nen)
      i_sl   = alloc_8('sl      ',nst,nst)
      i_ld   = alloc_4('ld      ',  1,nst)
c ... CSRC structure (Main.f:486-494)
      sia = 'ia      '
      sja = 'ja      '
      sau = 'au      '
      sal = 'al      '
      sad = 'ad      '
      call datastruct(ia(i_ix),ia(i_id),ia(i_inum),nnode,numel,nen,
     .                ndf,nst,neq,1,.false.,nad,nad1,i_ia,i_ja,i_au,
     .                i_al,i_ad,sia,sja,sau,sal,sad,ovlp)
      call coloredmesh(ia(i_ix),nnode,nnodev,numel,nenv,nen,numcolors,
     .                 i_colorg,i_elcolor)
      i_x0  = alloc_8('x0      ',1,neq)
      i_bst0= alloc_8('bstress0',1,neq)
      i_b0  = alloc_8('b0      ',1,neq)
      i_b   = alloc_8('b       ',1,neq)
      i_m   = alloc_8('m       ',1,neq)
      call azero(ia(i_x0),neq)
      call azero(ia(i_bst0),neq)
      call azero(ia(i_b0),neq)
      call azero(ia(i_b),neq)
      call azero(ia(i_m),neq)
c ... solvm (Main.f:718-790): nodal loads, element loads, then K and b
      call pload_mec(ia(i_id),ia(i_f),ia(i_u),ia(i_b0),ia(i_nload),
     .               nnode,ndf)
      nlit = 1
      call pform_mec(ia(i_ix),ia(i_eload),ia(i_ie),ia(i_e),ia(i_x),
     .        ia(i_id),ia(i_ia),ia(i_ja),ia(i_au),ia(i_al),ia(i_ad),
     .        ia(i_b0),ia(i_u),ia(i_tx0),ia(i_xl),ia(i_ul),ia(i_pl),
     .        ia(i_sl),ia(i_ld),ia(i_txnl),numel,nen,nenv,ndf,ndm,nst,
     .        neq,nad,nad1,.false.,.true.,.false.,1,4,1,nlit,
     .        ia(i_colorg),ia(i_elcolor),numcolors,.false.)
      call aequalb(ia(i_b),ia(i_b0),neq)
      call pform_mec(ia(i_ix),ia(i_eload),ia(i_ie),ia(i_e),ia(i_x),
     .        ia(i_id),ia(i_ia),ia(i_ja),ia(i_au),ia(i_al),ia(i_ad),
     .        ia(i_b),ia(i_u),ia(i_tx0),ia(i_xl),ia(i_ul),ia(i_pl),
     .        ia(i_sl),ia(i_ld),ia(i_txnl),numel,nen,nenv,ndf,ndm,nst,
     .        neq,nad,nad1,.true.,.true.,.false.,1,2,1,nlit,
     .        ia(i_colorg),ia(i_elcolor),numcolors,.false.)
c ... dumps of the set-up
      open(20,file=trim(outd)//'/meta.txt')
      write(20,*) nnode,numel,nen,ndf,ndm,neq,nad,nth
      call wi4(trim(outd)//'/inum.i4',ia(i_inum),nnode)
      call wi4(trim(outd)//'/id.i4',ia(i_id),ndf*nnode)
      call wi4(trim(outd)//'/ia.i4',ia(i_ia),neq+1)
      call wi4(trim(outd)//'/ja.i4',ia(i_ja),nad)
      call wf8(trim(outd)//'/ad.f8',ia(i_ad),neq)
      call wf8(trim(outd)//'/al.f8',ia(i_al),nad)
      call wf8(trim(outd)//'/b.f8',ia(i_b),neq)
      call wi4(trim(outd)//'/elcolor.i4',ia(i_elcolor),numel)
      call wi4(trim(outd)//'/colorg.i4',ia(i_colorg),2*numcolors)
c ... one product y = A v (Matvec.f:528), v(i) = 1 + mod(i,7)/8
      i_v = alloc_8('spmvv   ',1,neq)
      i_y = alloc_8('spmvy   ',1,neq)
      call fillv(ia(i_v),neq)
      call azero(ia(i_y),neq)
      call matvec_csrcsym1(neq,ia(i_ia),ia(i_ja),0,0,ia(i_ad),
     .     ia(i_al),ia(i_al),0,ia(i_v),ia(i_y),neqf1i,neqf2i,
     .     i_fmap,i_xf,i_rcvs,i_dspl,tt)
      call wf8(trim(outd)//'/spmv_v.f8',ia(i_v),neq)
      call wf8(trim(outd)//'/spmv_y.f8',ia(i_y),neq)
      i_y = dealloc('spmvy   ')
      i_v = dealloc('spmvv   ')
      if (iand(mode,64) .ne. 0) goto 900
c ... the solve, timed as soltime is (Main.f:819-827)
      open(10,file=trim(outd)//'/solver.log')
      tt = MPI_Wtime()
      call solv(neq,nad,ia(i_ia),ia(i_ja),ia(i_ad),ia(i_au),ia(i_al),
     .          ia(i_m),ia(i_b),ia(i_x0),solvtol,maxit,50,.false.,1,
     .          neqf1,neqf2,neq3,neq4,neq_dot,i_fmap,i_xf,i_rcvs,
     .          i_dspl)
      tt = MPI_Wtime() - tt
      write(20,*) tt
      call wf8(trim(outd)//'/x.f8',ia(i_x0),neq)
c ... second pass of the nonlinear loop: b0 - K u (Main.f:789-815)
      call update_mec(nnode,ndf,ia(i_id),ia(i_u),ia(i_x0))
      call aequalb(ia(i_b),ia(i_b0),neq)
      nlit = 2
      call pform_mec(ia(i_ix),ia(i_eload),ia(i_ie),ia(i_e),ia(i_x),
     .        ia(i_id),ia(i_ia),ia(i_ja),ia(i_au),ia(i_al),ia(i_ad),
     .        ia(i_b),ia(i_u),ia(i_tx0),ia(i_xl),ia(i_ul),ia(i_pl),
     .        ia(i_sl),ia(i_ld),ia(i_txnl),numel,nen,nenv,ndf,ndm,nst,
     .        neq,nad,nad1,.true.,.true.,.false.,1,2,1,nlit,
     .        ia(i_colorg),ia(i_elcolor),numcolors,.false.)
      call wf8(trim(outd)//'/resid.f8',ia(i_b),neq)
      call aequalb(ia(i_b),ia(i_b0),neq)
c ... GMRES(50) with its own signature (Iters.f:528)
      if (iand(mode,2) .ne. 0) then
        ngram = 50
        call aequalb(ia(i_m),ia(i_ad),neq)
        call pre_diag(ia(i_m),ia(i_m),neq,.false.)
        i_g = alloc_8('qg      ',neq,ngram+1)
        i_h = alloc_8('qh      ',ngram+1,ngram)
        i_y = alloc_8('qy      ',1,ngram)
        i_c = alloc_8('qc      ',1,ngram)
        i_s = alloc_8('qs      ',1,ngram)
        i_t = alloc_8('qe      ',1,ngram+1)
        call azero(ia(i_x0),neq)
        call gmres(neq,nad,ia(i_ia),ia(i_ja),ia(i_ad),ia(i_al),
     .       ia(i_al),ia(i_m),ia(i_b),ia(i_x0),ngram,ia(i_g),ia(i_h),
     .       ia(i_y),ia(i_c),ia(i_s),ia(i_t),solvtol,maxit,
     .       matvec_csrcsym1,dot_par,neq,my_id,neqf1i,neqf2i,neq_dot,
     .       i_fmap,i_xf,i_rcvs,i_dspl,.true.)
        call wf8(trim(outd)//'/x_gmres.f8',ia(i_x0),neq)
        i_t = dealloc('qe      ')
        i_s = dealloc('qs      ')
        i_c = dealloc('qc      ')
        i_y = dealloc('qy      ')
        i_h = dealloc('qh      ')
        i_g = dealloc('qg      ')
      endif
c ... BiCGSTAB with its own signature (Iters.f:1079)
      if (iand(mode,4) .ne. 0) then
        call aequalb(ia(i_m),ia(i_ad),neq)
        call pre_diag(ia(i_m),ia(i_m),neq,.false.)
        i_t  = alloc_8('qt      ',1,neq)
        i_v  = alloc_8('qv      ',1,neq)
        i_r  = alloc_8('qr      ',1,neq)
        i_p  = alloc_8('qp      ',1,neq)
        i_z  = alloc_8('qz      ',1,neq)
        i_r0 = alloc_8('qr0     ',1,neq)
        call azero(ia(i_x0),neq)
        call pbicgstab(neq,nad,ia(i_ia),ia(i_ja),ia(i_ad),ia(i_al),
     .       ia(i_al),ia(i_m),ia(i_b),ia(i_x0),ia(i_t),ia(i_v),ia(i_r),
     .       ia(i_p),ia(i_z),ia(i_r0),solvtol,maxit,matvec_csrcsym1,
     .       dot_par,my_id,neqf1i,neqf2i,neq_dot,i_fmap,i_xf,i_rcvs,
     .       i_dspl,.true.,.true.,.true.)
        call wf8(trim(outd)//'/x_bicgstab.f8',ia(i_x0),neq)
        i_r0 = dealloc('qr0     ')
        i_z  = dealloc('qz      ')
        i_p  = dealloc('qp      ')
        i_r  = dealloc('qr      ')
        i_v  = dealloc('qv      ')
        i_t  = dealloc('qt      ')
      endif
c ... 3x3 block-Jacobi set-up (Precond.f:163), dead in the driver
      if (iand(mode,8) .ne. 0) then
        param(1) = 0
        param(2) = 0
        param(3) = 6
        param(4) = 3
        mm = neq/3 + mod(neq,3)
        i_bd  = alloc_8('bdiag   ',6,mm)
        i_aux = alloc_8('baux    ',3,3)
        call block_precond(ia(i_ad),ia(i_al),ia(i_ia),ia(i_ja),
     .       ia(i_bd),neq,ia(i_aux),param)
        nbd = 6*int(mm)
        call wf8(trim(outd)//'/bdiag.f8',ia(i_bd),nbd)
c ... one application y = M^-1 v (Precond.f:419), v as for the product
        i_v = alloc_8('bdv     ',1,neq)
        i_y = alloc_8('bdy     ',1,neq)
        call fillv(ia(i_v),neq)
        call azero(ia(i_y),neq)
        call op_block_precond(ia(i_bd),ia(i_v),ia(i_y),ia(i_aux),param)
        call wf8(trim(outd)//'/bdiag_y.f8',ia(i_y),neq)
        i_y = dealloc('bdy     ')
        i_v = dealloc('bdv     ')
      endif
c ... 2x2 block-Jacobi set-up and one application (Precond.f:163, 419)
      if (iand(mode,256) .ne. 0) then
        param(1) = 0
        param(2) = 0
        param(3) = 3
        param(4) = 2
        mm = neq/2 + mod(neq,2)
        i_bd  = alloc_8('bdiag2  ',3,mm)
        i_aux = alloc_8('baux2   ',2,2)
        call block_precond(ia(i_ad),ia(i_al),ia(i_ia),ia(i_ja),
     .       ia(i_bd),neq,ia(i_aux),param)
        nbd = 3*int(mm)
        call wf8(trim(outd)//'/bdiag2.f8',ia(i_bd),nbd)
        i_v = alloc_8('bd2v    ',1,neq)
        i_y = alloc_8('bd2y    ',1,neq)
        call fillv(ia(i_v),neq)
        call azero(ia(i_y),neq)
        call op_block_precond(ia(i_bd),ia(i_v),ia(i_y),ia(i_aux),param)
        call wf8(trim(outd)//'/bdiag2_y.f8',ia(i_y),neq)
      endif
c ... nodal stresses and the result files (Main.f:930-963), after the
c     displacements of the solve went to u (update_mec above)
      if (iand(mode,32) .ne. 0) then
        ntn = 4
        if (ndm .eq. 3) ntn = 6
        i_tx = alloc_8('tx      ',ntn,nnode)
        i_ic = alloc_4('ic      ',1,nnode)
        call azero(ia(i_tx),nnode*ntn)
        call mzero(ia(i_ic),nnode)
        call tform_mec(ia(i_ix),ia(i_x),ia(i_e),ia(i_ie),ia(i_ic),
     .       ia(i_xl),ia(i_ul),ia(i_txnl),ia(i_u),ia(i_tx0),ia(i_tx),
     .       nnodev,nnode,numel,nenv,nen,ndm,ndf,nst,ntn,3,.false.,1)
        call wf8(trim(outd)//'/tx.f8',ia(i_tx),ntn*nnode)
        call wf8(trim(outd)//'/u.f8',ia(i_u),ndf*nnode)
        pflag = .false.
        pflag(2) = .true.
        pflag(3) = .true.
        prename = trim(outd)//'/res'
        call write_mesh_res_mec(ia(i_ix),ia(i_x),ia(i_u),ia(i_tx),
     .       nnode,numel,nen,ndm,ndf,ntn,prename,0,.false.,.true.,
     .       pflag,31)
        call write_mesh_res_mec(ia(i_ix),ia(i_x),ia(i_u),ia(i_tx),
     .       nnode,numel,nen,ndm,ndf,ntn,prename,0,.false.,.false.,
     .       pflag,31)
        prename = trim(outd)//'/resb'
        call write_mesh_res_mec(ia(i_ix),ia(i_x),ia(i_u),ia(i_tx),
     .       nnode,numel,nen,ndm,ndf,ntn,prename,0,.true.,.true.,
     .       pflag,31)
      endif
  900 continue
      close(10)
      close(20)
      call MPI_FINALIZE(ierr)
      end
c **********************************************************************
      subroutine wi4(fname,a,n)
      implicit none
      character*(*) fname
      integer n,a(n)
      open(30,file=fname,access='stream',form='unformatted',
     .     status='replace')
      write(30) a
      close(30)
      return
      end
c **********************************************************************
      subroutine wf8(fname,a,n)
      implicit none
      character*(*) fname
      integer n
      real*8 a(n)
      open(30,file=fname,access='stream',form='unformatted',
     .     status='replace')
      write(30) a
      close(30)
      return
      end
c **********************************************************************
      subroutine ri4(fname,a,n)
      implicit none
      character*(*) fname
      integer n,a(n)
      open(30,file=fname,access='stream',form='unformatted',
     .     status='old')
      read(30) a
      close(30)
      return
      end
c **********************************************************************
      subroutine rf8(fname,a,n)
      implicit none
      character*(*) fname
      integer n
      real*8 a(n)
      open(30,file=fname,access='stream',form='unformatted',
     .     status='old')
      read(30) a
      close(30)
      return
      end
c **********************************************************************
      subroutine fillv(v,n)
      implicit none
      integer n,i
      real*8 v(n)
      do i = 1, n
        v(i) = 1.d0 + dble(mod(i,7))*0.125d0
      enddo
      return
      end
