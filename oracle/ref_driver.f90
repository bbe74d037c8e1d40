!> Checker-side driver (test infrastructure only): exposes the reference's own
!> grid build, setBC/syncState/findEuler, findEulerJac and SPARSKIT ILU through
!> bind(c) entry points so tests can compare the C oracle against the compiled
!> reference.  It holds no algorithm of its own: every numeric step is a call
!> into the reference modules (modFOEuler, modEuler, modPolyFvGrid, modSparse)
!> built from /root/reference by oracle/Makefile.  The only lines restated here
!> are the iBC mapping and the dxdt=d/v division, copied in meaning from
!> src/foeuler/foeuler.f90:145-154 and :366-372.
module fosref
  use iso_c_binding
  use modFOEuler
  use modEuler
  implicit none
contains

  subroutine ref_grid_from_arrays(nN,pN,nE,sE,nNE,maxNNE,iNE,gidIn) bind(c,name='ref_grid_from_arrays')
    integer(C_INT),value::nN,nE,maxNNE
    real(C_DOUBLE),intent(in)::pN(3,nN)
    integer(C_INT),intent(in)::sE(nE),nNE(nE),iNE(maxNNE,nE),gidIn(nE)
    integer::i,j
    call grid%clear()
    call grid%init(nN,nE,maxNNE)
    grid%pN(:,:)=pN(:,:)
    do i=1,nE
      grid%sE(i)=sE(i)
      grid%nNE(i)=nNE(i)
      do j=1,nNE(i)
        grid%iNE(j,i)=iNE(j,i)+1
      end do
      grid%gid(i)=gidIn(i)
    end do
    call grid%up()
  end subroutine

  !> install already-built polyFvGrid arrays (0-based in, 1-based stored) and mark the grid up,
  !> so findEuler's grid%up() is a no-op; used to time the reference on large meshes
  subroutine ref_grid_set_fv(nN,nE,nC,nP,nPi,nPb,maxNF,maxNNE,pN,sE,nNE,iNE,gidIn,pE,v,iEP,aP,normP,pP,neib) &
  &  bind(c,name='ref_grid_set_fv')
    integer(C_INT),value::nN,nE,nC,nP,nPi,nPb,maxNF,maxNNE
    real(C_DOUBLE),intent(in)::pN(3,nN),pE(3,nE),v(nC),aP(nP),normP(3,nP),pP(3,nP)
    integer(C_INT),intent(in)::sE(nE),nNE(nE),iNE(maxNNE,nE),gidIn(nE),iEP(2,nP),neib(maxNF,nC)
    integer::i
    call grid%clear()
    call grid%init(nN,nE,maxNNE)
    grid%pN(:,:)=pN(:,:)
    do i=1,nE
      grid%sE(i)=sE(i)
      grid%nNE(i)=nNE(i)
      grid%iNE(:,i)=iNE(:,i)+1
      grid%gid(i)=gidIn(i)
      grid%p(:,i)=pE(:,i)
    end do
    grid%nC=nC
    grid%nF=nE-nC
    grid%nP=nP
    grid%nPi=nPi
    grid%nPb=nPb
    allocate(grid%v(nC),grid%iEP(2,nP),grid%neib(maxNF,nC),grid%aP(nP),grid%normP(3,nP),grid%pP(3,nP))
    do i=1,nC
      grid%v(i)=v(i)
      grid%neib(:,i)=neib(:,i)+1
    end do
    do i=1,nP
      grid%iEP(:,i)=iEP(:,i)+1
      grid%aP(i)=aP(i)
      grid%normP(:,i)=normP(:,i)
      grid%pP(:,i)=pP(:,i)
    end do
    grid%isUp=.true.
  end subroutine

  subroutine ref_grid_read_gmsh(path,n) bind(c,name='ref_grid_read_gmsh')
    use modFileIO
    integer(C_INT),value::n
    character(kind=C_CHAR),intent(in)::path(n)
    character(len=n)::fname
    integer::i
    do i=1,n
      fname(i:i)=path(i)
    end do
    call grid%clear()
    open(11,file=fname,action='read')
    call readGMSH(11,grid)
    close(11)
    call grid%up()
  end subroutine

  subroutine ref_grid_sizes(sz) bind(c,name='ref_grid_sizes')
    integer(C_INT),intent(out)::sz(10)
    sz(1)=grid%nN
    sz(2)=grid%nE
    sz(3)=grid%nC
    sz(4)=grid%nF
    sz(5)=grid%nP
    sz(6)=grid%nPi
    sz(7)=grid%nPb
    sz(8)=size(grid%neib,1)
    sz(9)=size(grid%iNE,1)
    sz(10)=0
  end subroutine

  subroutine ref_grid_get(iEP,aP,normP,pP,v,neib,pEl,sE,gidOut,iNE,pN) bind(c,name='ref_grid_get')
    integer(C_INT),intent(out)::iEP(2,grid%nP),neib(size(grid%neib,1),grid%nC),sE(grid%nE),gidOut(grid%nE)
    integer(C_INT),intent(out)::iNE(size(grid%iNE,1),grid%nE)
    real(C_DOUBLE),intent(out)::aP(grid%nP),normP(3,grid%nP),pP(3,grid%nP),v(grid%nC),pEl(3,grid%nE)
    real(C_DOUBLE),intent(out)::pN(3,grid%nN)
    iEP(:,:)=grid%iEP(:,1:grid%nP)-1
    aP(:)=grid%aP(1:grid%nP)
    normP(:,:)=grid%normP(:,1:grid%nP)
    pP(:,:)=grid%pP(:,1:grid%nP)
    v(:)=grid%v(1:grid%nC)
    neib(:,:)=grid%neib(:,:)-1
    pEl(:,:)=grid%p(:,:)
    sE(:)=grid%sE(:)
    gidOut(:)=grid%gid(:)
    iNE(:,:)=grid%iNE(:,:)-1
    pN(:,:)=grid%pN(:,:)
  end subroutine

  subroutine ref_setup_state(nCond,maxPar,condGid,condT,condP,rIn,gammIn) bind(c,name='ref_setup_state')
    integer(C_INT),value::nCond,maxPar
    integer(C_INT),intent(in)::condGid(nCond),condT(nCond)
    real(C_DOUBLE),intent(in)::condP(maxPar,nCond)
    real(C_DOUBLE),value::rIn,gammIn
    integer::i,j
    call bc%init(nCond,maxPar)
    bc%gid(:)=condGid(:)
    bc%t(:)=condT(:)
    bc%p(:,:)=condP(:,:)
    r=rIn
    gamm=gammIn
    if(allocated(rho)) deallocate(rho,rhou,rhoE,u,p,temp,c)
    if(allocated(iBC)) deallocate(iBC)
    if(allocated(dRho)) deallocate(dRho,dRhou,dRhoE)
    allocate(rho(grid%nE),rhou(3,grid%nE),rhoE(grid%nE),u(3,grid%nE),p(grid%nE),temp(grid%nE),c(grid%nE))
    rho(:)=1d0
    rhou(:,:)=0d0
    rhoE(:)=1d0
    allocate(iBC(grid%nE))
    iBC(:)=0
    do i=1,grid%nE
      do j=1,size(bc%gid)
        if(grid%gid(i)==bc%gid(j))then
          iBC(i)=j
          exit
        end if
      end do
    end do
    allocate(dRho(grid%nC),dRhou(3,grid%nC),dRhoE(grid%nC))
  end subroutine

  !> setBC + syncState + findEuler + division by v, i.e. the body of foeuler's rhs
  subroutine ref_rhs(x,dxdt) bind(c,name='ref_rhs')
    real(C_DOUBLE),intent(in)::x(5*grid%nC)
    real(C_DOUBLE),intent(out)::dxdt(5*grid%nC)
    integer::i
    call setBC(x)
    call syncState(x)
    call findEuler(grid,rho,rhou,rhoE,p,gamm,dRho,dRhou,dRhoE)
    do i=1,grid%nC
      dxdt(i)=dRho(i)/grid%v(i)
      dxdt(grid%nC+i)=dRhou(1,i)/grid%v(i)
      dxdt(2*grid%nC+i)=dRhou(2,i)/grid%v(i)
      dxdt(3*grid%nC+i)=dRhou(3,i)/grid%v(i)
      dxdt(4*grid%nC+i)=dRhoE(i)/grid%v(i)
    end do
  end subroutine

  !> element state after the last ref_rhs (ghosts included)
  subroutine ref_get_state(rhoO,rhouO,rhoEO,pO) bind(c,name='ref_get_state')
    real(C_DOUBLE),intent(out)::rhoO(grid%nE),rhouO(3,grid%nE),rhoEO(grid%nE),pO(grid%nE)
    rhoO(:)=rho(:)
    rhouO(:,:)=rhou(:,:)
    rhoEO(:)=rhoE(:)
    pO(:)=p(:)
  end subroutine

  subroutine ref_find_euler(rhoI,rhouI,rhoEI,pI,gammI,dR,dRu,dRE) bind(c,name='ref_find_euler')
    real(C_DOUBLE),intent(in)::rhoI(grid%nE),rhouI(3,grid%nE),rhoEI(grid%nE),pI(grid%nE)
    real(C_DOUBLE),value::gammI
    real(C_DOUBLE),intent(out)::dR(grid%nC),dRu(3,grid%nC),dRE(grid%nC)
    double precision,allocatable::a(:),b(:,:),e(:)
    call findEuler(grid,rhoI,rhouI,rhoEI,pI,gammI,a,b,e)
    dR(:)=a(:)
    dRu(:,:)=b(:,:)
    dRE(:)=e(:)
  end subroutine

  subroutine ref_find_euler_jac(rhoI,rhouI,rhoEI,pI,gammI,jac) bind(c,name='ref_find_euler_jac')
    real(C_DOUBLE),intent(in)::rhoI(grid%nE),rhouI(3,grid%nE),rhoEI(grid%nE),pI(grid%nE)
    real(C_DOUBLE),value::gammI
    real(C_DOUBLE),intent(out)::jac(5,5*size(grid%neib,1)+5,grid%nC)
    double precision,allocatable::jp(:,:,:),jc(:,:,:)
    call findEulerJac(grid,rhoI,rhouI,rhoEI,pI,gammI,jp,jc)
    jac(:,:,:)=jc(:,:,:)
  end subroutine

  !> ILU wrapper of the reference (sparse.f90:334-392): init, setCOO, fact, solve
  subroutine ref_ilu_solve(n,nnz,iA,jA,A,lFill,droptol,rhs,x) bind(c,name='ref_ilu_solve')
    use modSparse
    integer(C_INT),value::n,nnz,lFill
    integer(C_INT),intent(in)::iA(nnz),jA(nnz)
    real(C_DOUBLE),intent(in)::A(nnz),rhs(n)
    real(C_DOUBLE),value::droptol
    real(C_DOUBLE),intent(out)::x(n)
    type(ILU)::pc
    integer::ia1(nnz),ja1(nnz)
    ia1(:)=iA(:)
    ja1(:)=jA(:)
    call pc%init(n,nnz,lFill)
    call pc%setCOO(ia1,ja1,A,nnz)
    call pc%fact(droptol)
    call pc%solve(rhs,x)
    call pc%clear()
  end subroutine

  !> linEq%mulVec (sparse.f90:250-256) after setCOO
  subroutine ref_mulvec(n,nnz,iA,jA,A,xIn,y) bind(c,name='ref_mulvec')
    use modSparse
    integer(C_INT),value::n,nnz
    integer(C_INT),intent(in)::iA(nnz),jA(nnz)
    real(C_DOUBLE),intent(in)::A(nnz),xIn(n)
    real(C_DOUBLE),intent(out)::y(n)
    type(linEq)::le
    integer::ia1(nnz),ja1(nnz)
    ia1(:)=iA(:)
    ja1(:)=jA(:)
    call le%initLinEq(n,nnz)
    call le%setCOO(ia1,ja1,A,nnz)
    call le%mulVec(xIn,y)
    call le%clear()
  end subroutine

  !> SPARSKIT ilu0 + lusol on a 1-based CSR matrix (ilut.f:1793-1917, :2272-2321)
  subroutine ref_ilu0_solve(n,ia,ja,a,rhs,x,ierr) bind(c,name='ref_ilu0_solve')
    integer(C_INT),value::n
    integer(C_INT),intent(in)::ia(n+1),ja(*)
    real(C_DOUBLE),intent(in)::a(*),rhs(n)
    real(C_DOUBLE),intent(out)::x(n)
    integer(C_INT),intent(out)::ierr
    double precision,allocatable::alu(:)
    integer,allocatable::jlu(:),ju(:),iw(:)
    integer::nnz,ie
    external::ilu0,lusol
    nnz=ia(n+1)-1
    allocate(alu(nnz+n+2),jlu(nnz+n+2),ju(n),iw(n))
    call ilu0(n,a,ja,ia,alu,jlu,ju,iw,ie)
    ierr=ie
    if(ie==0) call lusol(n,rhs,x,alu,jlu,ju)
    deallocate(alu,jlu,ju,iw)
  end subroutine

end module
